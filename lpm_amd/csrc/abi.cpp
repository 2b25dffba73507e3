// C-ABI entry points: contexts, error reporting, meshes, direct sums and
// kernel timing.  Every entry point catches all exceptions and converts them
// to an lpm_status plus a message (LPM_REQUIRE in the reference throws
// std::runtime_error, src/lpm_assert.hpp:22-30,54; the C++ wrapper classes
// rethrow).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>

#include "internal.hpp"
#include "mesh.hpp"

using namespace lpm_amd;

namespace {
thread_local std::string g_thread_error = "no error";

template <typename F>
int guarded(lpm_ctx* ctx, F&& f) {
  try {
    f();
    return LPM_OK;
  } catch (const Error& e) {
    set_error(ctx, e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error(ctx, "host allocation failed");
    return LPM_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error(ctx, e.what());
    return LPM_ERR_INTERNAL;
  } catch (...) {
    set_error(ctx, "unknown exception");
    return LPM_ERR_INTERNAL;
  }
}
}  // namespace

struct lpm_mesh {
  std::unique_ptr<SphereTreeMesh> m;
};

namespace lpm_amd {

void set_error(lpm_ctx* ctx, const std::string& msg) {
  if (ctx)
    ctx->last_error = msg;
  else
    g_thread_error = msg;
}

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

void order_stream(lpm_ctx* ctx, hipStream_t s) {
  if (ctx->has_last_stream && ctx->last_stream == s) return;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  LPM_HIP_CHECK(hipStreamIsCapturing(s, &st));
  if (st != hipStreamCaptureStatusNone) return;
  if (ctx->has_last_stream) {
    if (!ctx->order_event)
      LPM_HIP_CHECK(hipEventCreateWithFlags(&ctx->order_event,
                                            hipEventDisableTiming));
    LPM_HIP_CHECK(hipEventRecord(ctx->order_event, ctx->last_stream));
    LPM_HIP_CHECK(hipStreamWaitEvent(s, ctx->order_event, 0));
  }
  ctx->last_stream = s;
  ctx->has_last_stream = true;
}

std::atomic<uint64_t> g_workspace_epoch{0};

DevBuf::~DevBuf() {
  if (ptr) (void)hipFree(ptr);
}

void* DevBuf::ensure(size_t nbytes) {
  if (nbytes <= bytes && ptr) return ptr;
  if (ptr) {
    LPM_HIP_CHECK(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
  }
  const size_t want = std::max<size_t>(nbytes, 256);
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    throw Error(LPM_ERR_NOMEM, std::string("hipMalloc of ") +
                                   std::to_string(want) +
                                   " bytes failed: " + hipGetErrorString(e));
  }
  bytes = want;
  g_workspace_epoch.fetch_add(1);
  return ptr;
}

// Size bucket: floor(log2(total targets)); tables are per device arch.
std::string tune_key(const lpm_ctx* ctx, int family, int64_t n_targets) {
  int b = 0;
  while ((int64_t(1) << (b + 1)) <= n_targets && b < 62) ++b;
  return ctx->arch + "/" + std::to_string(family) + "/" + std::to_string(b);
}

int pick_variant(lpm_ctx* ctx, int family, int64_t n_targets) {
  if (ctx->forced[family] >= 0) return ctx->forced[family];
  auto it = ctx->tuned.find(tune_key(ctx, family, n_targets));
  if (it != ctx->tuned.end() && it->second >= 0 &&
      it->second < variant_count(family))
    return it->second;
  return default_variant(family);
}

TimedLaunch::TimedLaunch(lpm_ctx* c, int k, hipStream_t s, double p)
    : ctx(c), kid(k), stream(s), pairs(p) {
  // kid < 0: an untimed launch
  if (kid < 0 || !ctx->timing) return;
  LPM_HIP_CHECK(hipEventCreate(&e0));
  LPM_HIP_CHECK(hipEventCreate(&e1));
  LPM_HIP_CHECK(hipEventRecord(e0, stream));
}

TimedLaunch::~TimedLaunch() noexcept(false) {
  if (!e0) return;
  if (hipEventRecord(e1, stream) != hipSuccess) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return;
  }
  auto& r = ctx->timing_rec[kid];
  r.pending.emplace_back(e0, e1);
  r.launches += 1;
  r.pairs += pairs;
}

}  // namespace lpm_amd

extern "C" {

int lpm_abi_version(void) { return LPM_ABI_VERSION; }

const char* lpm_last_error(const lpm_ctx* ctx) {
  return ctx ? ctx->last_error.c_str() : g_thread_error.c_str();
}

int lpm_ctx_create(int device, lpm_ctx** out) {
  return guarded(nullptr, [&] {
    LPM_ARG_CHECK(out != nullptr, "lpm_ctx_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    LPM_HIP_CHECK(hipGetDeviceCount(&ndev));
    LPM_ARG_CHECK(device >= 0 && device < ndev,
                  "lpm_ctx_create: device index out of range");
    LPM_HIP_CHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    LPM_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    auto ctx = std::make_unique<lpm_ctx>();
    ctx->device = device;
    ctx->num_cus = prop.multiProcessorCount;
    ctx->arch = prop.gcnArchName;
    ctx->last_error = "no error";
    // the device's copy of the allocation-free scratch array (once per device)
    noalloc_setup(device);
    *out = ctx.release();
  });
}

int lpm_ctx_destroy(lpm_ctx* ctx) {
  if (!ctx) return LPM_OK;
  (void)hipSetDevice(ctx->device);
  // captured graphs and workspaces may still be in use by queued work
  (void)hipDeviceSynchronize();
  if (ctx->order_event) (void)hipEventDestroy(ctx->order_event);
  for (auto& r : ctx->timing_rec)
    for (auto& ev : r.pending) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
  delete ctx;
  return LPM_OK;
}

int lpm_stream_synchronize(lpm_ctx* ctx, void* stream) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx, "null context");
    LPM_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  });
}

int lpm_graph_cache_size(const lpm_ctx* ctx, int* n) {
  return guarded(nullptr, [&] {
    LPM_ARG_CHECK(ctx && n, "null argument");
    *n = static_cast<int>(ctx->graphs.size());
  });
}

// ------------------------------------------------------------------ mesh
int lpm_mesh_create(int seed, int depth, lpm_mesh** out) {
  return guarded(nullptr, [&] {
    LPM_ARG_CHECK(out != nullptr, "lpm_mesh_create: out is NULL");
    LPM_ARG_CHECK(seed == LPM_SEED_ICOS_TRI || seed == LPM_SEED_CUBED_SPHERE,
                  "lpm_mesh_create: unknown seed");
    LPM_ARG_CHECK(depth >= 0 && depth <= 10,
                  "lpm_mesh_create: depth must be in [0, 10]");
    auto m = std::make_unique<lpm_mesh>();
    m->m = std::make_unique<SphereTreeMesh>(static_cast<SeedKind>(seed), depth);
    *out = m.release();
  });
}

int lpm_mesh_destroy(lpm_mesh* mesh) {
  delete mesh;
  return LPM_OK;
}

int lpm_mesh_sizes(const lpm_mesh* mesh, int64_t* nv, int64_t* nf,
                   int64_t* nl, int64_t* ne) {
  return guarded(nullptr, [&] {
    LPM_ARG_CHECK(mesh && mesh->m, "lpm_mesh_sizes: null mesh");
    if (nv) *nv = mesh->m->n_vertices();
    if (nf) *nf = mesh->m->n_faces();
    if (nl) *nl = mesh->m->n_leaves();
    if (ne) *ne = mesh->m->n_edges();
  });
}

int lpm_mesh_max_allocations(int seed, int depth, int64_t* nv, int64_t* ne,
                             int64_t* nf) {
  return guarded(nullptr, [&] {
    LPM_ARG_CHECK(seed == LPM_SEED_ICOS_TRI || seed == LPM_SEED_CUBED_SPHERE,
                  "unknown seed");
    LPM_ARG_CHECK(depth >= 0 && depth <= 12, "depth out of range");
    int64_t a, b, c;
    SeedCounts::max_allocations(static_cast<SeedKind>(seed), depth, a, b, c);
    if (nv) *nv = a;
    if (ne) *ne = b;
    if (nf) *nf = c;
  });
}

int lpm_mesh_copy_vertices(const lpm_mesh* mesh, double* xyz) {
  return guarded(nullptr, [&] {
    LPM_ARG_CHECK(mesh && mesh->m, "null mesh");
    const auto& m = *mesh->m;
    const int64_t n = m.n_vertices();
    if (xyz) {
      std::memcpy(xyz, m.vx().data(), sizeof(double) * n);
      std::memcpy(xyz + n, m.vy().data(), sizeof(double) * n);
      std::memcpy(xyz + 2 * n, m.vz().data(), sizeof(double) * n);
    }
  });
}

int lpm_mesh_copy_faces(const lpm_mesh* mesh, double* xyz, double* area,
                        uint8_t* mask, int32_t* parent) {
  return guarded(nullptr, [&] {
    LPM_ARG_CHECK(mesh && mesh->m, "null mesh");
    const auto& m = *mesh->m;
    const int64_t n = m.n_faces();
    if (xyz) {
      std::memcpy(xyz, m.fx().data(), sizeof(double) * n);
      std::memcpy(xyz + n, m.fy().data(), sizeof(double) * n);
      std::memcpy(xyz + 2 * n, m.fz().data(), sizeof(double) * n);
    }
    if (area) std::memcpy(area, m.farea().data(), sizeof(double) * n);
    if (mask) std::memcpy(mask, m.fmask().data(), n);
    if (parent) std::memcpy(parent, m.fparent().data(), sizeof(int32_t) * n);
  });
}

int lpm_mesh_copy_face_vertices(const lpm_mesh* mesh, int32_t* verts,
                                int* nfv) {
  return guarded(nullptr, [&] {
    LPM_ARG_CHECK(mesh && mesh->m, "null mesh");
    const auto& m = *mesh->m;
    if (nfv) *nfv = m.verts_per_face();
    if (verts)
      std::memcpy(verts, m.fverts().data(),
                  sizeof(int32_t) * m.fverts().size());
  });
}

int lpm_mesh_copy_face_levels(const lpm_mesh* mesh, int32_t* level) {
  return guarded(nullptr, [&] {
    LPM_ARG_CHECK(mesh && mesh->m, "null mesh");
    const auto& m = *mesh->m;
    if (level)
      std::memcpy(level, m.flevel().data(), sizeof(int32_t) * m.n_faces());
  });
}

int lpm_mesh_divide_flagged_faces(lpm_mesh* mesh, const uint8_t* flags,
                                  int64_t n_flags, int amr_limit,
                                  int64_t max_faces, int64_t* n_divided,
                                  int* limit_reached) {
  return guarded(nullptr, [&] {
    LPM_ARG_CHECK(mesh && mesh->m, "null mesh");
    auto& m = *mesh->m;
    LPM_ARG_CHECK(n_flags == m.n_faces(),
                  "lpm_mesh_divide_flagged_faces: n_flags must equal n_faces");
    LPM_ARG_CHECK(flags || n_flags == 0, "flags is NULL");
    LPM_ARG_CHECK(amr_limit >= 0, "amr_limit must be >= 0");
    int64_t flagged = 0;
    for (int64_t i = 0; i < n_flags; ++i) flagged += flags[i] ? 1 : 0;
    if (limit_reached) *limit_reached = 0;
    if (max_faces > 0 && flagged > (max_faces - m.n_faces()) / 4) {
      if (n_divided) *n_divided = -1;
      return;
    }
    bool lim = false;
    const int64_t c = m.divide_flagged(flags, m.depth() + amr_limit, &lim);
    if (n_divided) *n_divided = c;
    if (limit_reached) *limit_reached = lim ? 1 : 0;
  });
}

// ------------------------------------------------------------ BVE sums
static void check_crd(const double* p, int64_t es, int64_t cs, int64_t n,
                      const char* what) {
  if (n == 0) return;
  LPM_ARG_CHECK(p != nullptr, std::string(what) + " is NULL");
  LPM_ARG_CHECK(es > 0 && cs > 0, std::string(what) + ": strides must be > 0");
}

int lpm_bve_velocity(lpm_ctx* ctx, const double* tgt_x, int64_t tgt_es,
                     int64_t tgt_cs, int64_t n_tgt, const double* src_x,
                     int64_t src_es, int64_t src_cs, const double* src_zeta,
                     const double* src_area, const uint8_t* src_mask,
                     int64_t n_src, int collocated, double eps, int precision,
                     unsigned flags, double* out_u, int64_t out_es,
                     int64_t out_cs, void* stream) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx, "null context");
    LPM_ARG_CHECK(n_tgt >= 0 && n_src >= 0, "negative size");
    LPM_ARG_CHECK(precision == LPM_FP64 || precision == LPM_MIXED,
                  "unknown precision");
    LPM_ARG_CHECK(eps >= 0, "eps must be >= 0");
    check_crd(tgt_x, tgt_es, tgt_cs, n_tgt, "tgt_x");
    check_crd(src_x, src_es, src_cs, n_src, "src_x");
    check_crd(out_u, out_es, out_cs, n_tgt, "out_u");
    LPM_ARG_CHECK(n_src == 0 || (src_zeta && src_area),
                  "src_zeta/src_area are NULL");
    if (collocated)
      LPM_ARG_CHECK(n_tgt == n_src,
                    "collocated: targets must be the sources (n_tgt == n_src)");
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const RankMap map = build_rank(ctx, src_mask, n_src, s);
    PackedSources ps = pack_sources(ctx, CrdIn{src_x, src_es, src_cs},
                                    src_zeta, src_area, map, precision, s);
    TargetSeg t;
    if (collocated) {
      t = face_targets(CrdIn{tgt_x, tgt_es, tgt_cs}, map, src_mask,
                       (flags & LPM_FLAG_SKIP_MASKED_TARGETS) != 0,
                       CrdOut{out_u, out_es, out_cs}, nullptr);
    } else {
      t.x = CrdIn{tgt_x, tgt_es, tgt_cs};
      t.n = n_tgt;
      t.out = CrdOut{out_u, out_es, out_cs};
    }
    launch_bve_velocity(ctx, ps, &t, 1, eps * eps, precision, s);
    if (flags & LPM_FLAG_CHECK_FINITE)
      check_finite(ctx, out_u, n_tgt, out_es, out_cs, 3, s);
  });
}

int lpm_bve_stream_fn(lpm_ctx* ctx, const double* tgt_x, int64_t tgt_es,
                      int64_t tgt_cs, int64_t n_tgt, const double* src_x,
                      int64_t src_es, int64_t src_cs, const double* src_zeta,
                      const double* src_area, const uint8_t* src_mask,
                      int64_t n_src, int collocated, int precision,
                      unsigned flags, double* out_psi, void* stream) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx, "null context");
    LPM_ARG_CHECK(n_tgt >= 0 && n_src >= 0, "negative size");
    LPM_ARG_CHECK(precision == LPM_FP64 || precision == LPM_MIXED,
                  "unknown precision");
    check_crd(tgt_x, tgt_es, tgt_cs, n_tgt, "tgt_x");
    check_crd(src_x, src_es, src_cs, n_src, "src_x");
    LPM_ARG_CHECK(n_tgt == 0 || out_psi, "out_psi is NULL");
    LPM_ARG_CHECK(n_src == 0 || (src_zeta && src_area),
                  "src_zeta/src_area are NULL");
    if (collocated)
      LPM_ARG_CHECK(n_tgt == n_src,
                    "collocated: targets must be the sources (n_tgt == n_src)");
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const RankMap map = build_rank(ctx, src_mask, n_src, s);
    // the stream function is evaluated in FP64 for both precisions
    PackedSources ps = pack_sources(ctx, CrdIn{src_x, src_es, src_cs},
                                    src_zeta, src_area, map, LPM_FP64, s);
    TargetSeg t;
    if (collocated) {
      t = face_targets(CrdIn{tgt_x, tgt_es, tgt_cs}, map, src_mask,
                       (flags & LPM_FLAG_SKIP_MASKED_TARGETS) != 0,
                       CrdOut{nullptr, 0, 0}, out_psi);
    } else {
      t.x = CrdIn{tgt_x, tgt_es, tgt_cs};
      t.n = n_tgt;
      t.out_scalar = out_psi;
    }
    launch_bve_stream_fn(ctx, ps, &t, 1, s);
    if (flags & LPM_FLAG_CHECK_FINITE)
      check_finite(ctx, out_psi, n_tgt, 1, 1, 1, s);
  });
}

// ----------------------------------------------------------- variants
int lpm_variant_count(int kernel_id) { return variant_count(kernel_id); }

int lpm_variant_default(int kernel_id) {
  return variant_count(kernel_id) > 0 ? default_variant(kernel_id) : -1;
}

const char* lpm_variant_name(int kernel_id, int variant) {
  const char* n = variant_name(kernel_id, variant);
  return n ? n : "";
}

int lpm_set_variant(lpm_ctx* ctx, int kernel_id, int variant) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx, "null context");
    LPM_ARG_CHECK(kernel_id >= 0 && kernel_id < KID_COUNT, "bad kernel id");
    LPM_ARG_CHECK(variant >= -1 && variant < variant_count(kernel_id),
                  "variant out of range");
    ctx->forced[kernel_id] = variant;
    ++ctx->choice_serial;
  });
}

int lpm_tune_save(lpm_ctx* ctx, const char* path) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx && path, "null argument");
    FILE* f = std::fopen(path, "w");
    LPM_ARG_CHECK(f != nullptr, std::string("cannot write ") + path);
    for (const auto& kv : ctx->tuned)
      std::fprintf(f, "%s %d\n", kv.first.c_str(), kv.second);
    std::fclose(f);
  });
}

int lpm_tune_load(lpm_ctx* ctx, const char* path) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx && path, "null argument");
    FILE* f = std::fopen(path, "r");
    LPM_ARG_CHECK(f != nullptr, std::string("cannot read ") + path);
    char key[256];
    int v = 0;
    while (std::fscanf(f, "%255s %d", key, &v) == 2) ctx->tuned[key] = v;
    ++ctx->choice_serial;
    std::fclose(f);
  });
}

// -------------------------------------------------------------- timing
int lpm_timing_enable(lpm_ctx* ctx, int enable) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx, "null context");
    for (auto& r : ctx->timing_rec) {
      for (auto& ev : r.pending) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
      }
      r = TimingRec{};
    }
    ctx->timing = enable != 0;
  });
}

int lpm_timing_query(lpm_ctx* ctx, int kernel_id, double* total_ms,
                     int64_t* launches, double* pairs) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx, "null context");
    LPM_ARG_CHECK(kernel_id >= 0 && kernel_id < KID_COUNT, "bad kernel id");
    auto& r = ctx->timing_rec[kernel_id];
    for (auto& ev : r.pending) {
      LPM_HIP_CHECK(hipEventSynchronize(ev.second));
      float ms = 0;
      LPM_HIP_CHECK(hipEventElapsedTime(&ms, ev.first, ev.second));
      r.total_ms += ms;
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
    r.pending.clear();
    if (total_ms) *total_ms = r.total_ms;
    if (launches) *launches = r.launches;
    if (pairs) *pairs = r.pairs;
  });
}

}  // extern "C"
