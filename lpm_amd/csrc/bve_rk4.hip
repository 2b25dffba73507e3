// BVE RK4 stepper on the device: fused O(N) stage kernels around the direct
// sums, following BVERK4::advance_timestep (src/lpm_bve_rk4_impl.hpp:55-167)
// stage by stage.  The reference's per-stage KokkosBlas::scal/update calls
// and BVEVorticityTendency launches (30 O(N) launches per step) become one
// stage kernel per particle kind and stage.  Optionally the whole step is
// captured into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <functional>
#include <sstream>

#include "internal.hpp"

struct lpm_ctx::GraphEntry {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t epoch = 0;
  uint64_t last_use = 0;          // LRU stamp
  hipStream_t stream = nullptr;   // stream of the last replay
  ~GraphEntry() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace lpm_amd {

namespace {

// Stage storage of one particle kind, SoA with leading dimension n:
//   k[s]  : dt*u of stage s (3n), s = 0..3
//   zk[s] : vorticity increment of stage s (n)
//   xw    : work positions (3n), zw : work vorticity (n)
struct StageBufs {
  double* k[4];
  double* zk[4];
  double* xw;
  double* zw;
  int64_t n;
};

StageBufs carve(DevBuf& buf, int64_t n) {
  const int64_t nn = std::max<int64_t>(n, 1);
  double* p = static_cast<double*>(buf.ensure(sizeof(double) * 20 * nn));
  StageBufs b;
  for (int s = 0; s < 4; ++s) {
    b.k[s] = p + (4 * s) * nn;
    b.zk[s] = p + (4 * s + 3) * nn;
  }
  b.xw = p + 16 * nn;
  b.zw = p + 19 * nn;
  b.n = n;
  return b;
}

struct StageArgs {
  const double* x;  // state positions (strided)
  int64_t xes, xcs;
  const double* zeta;
  const double* u;  // velocity just evaluated (strided)
  int64_t ues, ucs;
  const uint8_t* skip;
  double* k;   // this stage's dt*u (SoA, ld = n)
  double* zk;  // this stage's vorticity increment
  double* xw;  // next work positions (SoA), nullptr after stage 4
  double* zw;
  int64_t n;
  double dt, Omega, a;
  int renorm;
};

// x_k = dt u; zeta_k = -2 Omega u_z dt (BVEVorticityTendency);
// xwork = x + a x_k, zetawork = zeta + a zeta_k (KokkosBlas::update).
__global__ __launch_bounds__(256) void rk_stage_kernel(StageArgs A) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  if (A.skip && A.skip[i]) return;
  const double* u = A.u + i * A.ues;
  const double u0 = u[0], u1 = u[A.ucs], u2 = u[2 * A.ucs];
  const double k0 = A.dt * u0, k1 = A.dt * u1, k2 = A.dt * u2;
  const double zk = -2.0 * A.Omega * u2 * A.dt;
  const int64_t n = A.n;
  A.k[i] = k0;
  A.k[i + n] = k1;
  A.k[i + 2 * n] = k2;
  A.zk[i] = zk;
  if (A.xw) {
    const double* x = A.x + i * A.xes;
    double w0 = x[0] + A.a * k0;
    double w1 = x[A.xcs] + A.a * k1;
    double w2 = x[2 * A.xcs] + A.a * k2;
    if (A.renorm) {
      const double r = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2);
      w0 *= r;
      w1 *= r;
      w2 *= r;
    }
    A.xw[i] = w0;
    A.xw[i + n] = w1;
    A.xw[i + 2 * n] = w2;
    A.zw[i] = A.zeta[i] + A.a * zk;
  }
}

struct UpdateArgs {
  double* x;
  int64_t xes, xcs;
  double* zeta;
  const uint8_t* skip;
  const double* k[4];
  const double* zk[4];  // already permuted for compat mode
  int64_t n;
  int renorm;
};

// BVERK4Update: x += (k1 + k4)/6 + (k2 + k3)/3, same for zeta.
__global__ __launch_bounds__(256) void rk_update_kernel(UpdateArgs A) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  if (A.skip && A.skip[i]) return;
  constexpr double sixth = 1.0 / 6.0, third = 1.0 / 3.0;
  const int64_t n = A.n;
  double* x = A.x + i * A.xes;
  double v[3];
  for (int c = 0; c < 3; ++c) {
    const int64_t o = i + c * n;
    v[c] = x[c * A.xcs] + (sixth * (A.k[0][o] + A.k[3][o]) +
                           third * (A.k[1][o] + A.k[2][o]));
  }
  if (A.renorm) {
    const double r = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] *= r;
    v[1] *= r;
    v[2] *= r;
  }
  for (int c = 0; c < 3; ++c) x[c * A.xcs] = v[c];
  A.zeta[i] += sixth * (A.zk[0][i] + A.zk[3][i]) +
               third * (A.zk[1][i] + A.zk[2][i]);
}

inline int nblk(int64_t n) { return static_cast<int>((n + 255) / 256); }

// Velocity at vertex positions vx and face positions fx with face vorticity
// fzeta, written to f.vert_u / f.face_u.
using EvalFn =
    std::function<void(const RankMap&, CrdIn vx, CrdIn fx, const double* fz)>;

void velocity_eval(lpm_ctx* ctx, const lpm_bve_fields& f, CrdIn vx, CrdIn fx,
                   const double* fzeta, const RankMap& map, int precision,
                   unsigned flags, hipStream_t s) {
  PackedSources ps =
      pack_sources(ctx, fx, fzeta, f.face_area, map, precision, s);
  TargetSeg segs[2];
  segs[0].x = vx;
  segs[0].n = f.n_verts;
  segs[0].out = CrdOut{f.vert_u, f.vert_u_es, f.vert_u_cs};
  // vertex slots in the spatial order of the state positions (the stage
  // positions of a step are close to them)
  if (precision == LPM_MIXED)
    segs[0].index = target_order(ctx, CrdIn{f.vert_x, f.vert_x_es, f.vert_x_cs},
                                 f.n_verts, nullptr, 1, s);
  segs[1] = face_targets(fx, map, f.face_mask,
                         (flags & LPM_FLAG_SKIP_MASKED_TARGETS) != 0,
                         CrdOut{f.face_u, f.face_u_es, f.face_u_cs}, nullptr);
  launch_bve_velocity(ctx, ps, segs, 2, 0.0, precision, s, -1, true, true);
}

CrdIn vert_crd(const lpm_bve_fields& f) {
  return CrdIn{f.vert_x, f.vert_x_es, f.vert_x_cs};
}
CrdIn face_crd(const lpm_bve_fields& f) {
  return CrdIn{f.face_x, f.face_x_es, f.face_x_cs};
}

void issue_step(lpm_ctx* ctx, const lpm_bve_fields& f, double dt,
                double Omega, unsigned flags, StageBufs& V, StageBufs& F,
                const EvalFn& eval, hipStream_t s) {
  const int renorm = (flags & LPM_RK4_RENORMALIZE) ? 1 : 0;
  const uint8_t* fskip =
      (flags & LPM_FLAG_SKIP_MASKED_TARGETS) ? f.face_mask : nullptr;
  const RankMap map = build_rank(ctx, f.face_mask, f.n_faces, s);
  const double coef[3] = {0.5, 0.5, 1.0};

  for (int stage = 0; stage < 4; ++stage) {
    if (stage > 0)
      eval(map, CrdIn{V.xw, 1, V.n}, CrdIn{F.xw, 1, F.n}, F.zw);
    StageArgs a{};
    a.dt = dt;
    a.Omega = Omega;
    a.a = stage < 3 ? coef[stage] : 0.0;
    a.renorm = renorm;
    if (f.n_verts > 0) {
      a.x = f.vert_x;
      a.xes = f.vert_x_es;
      a.xcs = f.vert_x_cs;
      a.zeta = f.vert_zeta;
      a.u = f.vert_u;
      a.ues = f.vert_u_es;
      a.ucs = f.vert_u_cs;
      a.skip = nullptr;
      a.k = V.k[stage];
      a.zk = V.zk[stage];
      a.xw = stage < 3 ? V.xw : nullptr;
      a.zw = V.zw;
      a.n = f.n_verts;
      rk_stage_kernel<<<nblk(a.n), 256, 0, s>>>(a);
      LPM_HIP_CHECK(hipGetLastError());
    }
    a.x = f.face_x;
    a.xes = f.face_x_es;
    a.xcs = f.face_x_cs;
    a.zeta = f.face_zeta;
    a.u = f.face_u;
    a.ues = f.face_u_es;
    a.ucs = f.face_u_cs;
    a.skip = fskip;
    a.k = F.k[stage];
    a.zk = F.zk[stage];
    a.xw = stage < 3 ? F.xw : nullptr;
    a.zw = F.zw;
    a.n = f.n_faces;
    rk_stage_kernel<<<nblk(a.n), 256, 0, s>>>(a);
    LPM_HIP_CHECK(hipGetLastError());
  }

  UpdateArgs u{};
  u.renorm = renorm;
  if (f.n_verts > 0) {
    u.x = f.vert_x;
    u.xes = f.vert_x_es;
    u.xcs = f.vert_x_cs;
    u.zeta = f.vert_zeta;
    u.skip = nullptr;
    for (int k = 0; k < 4; ++k) {
      u.k[k] = V.k[k];
      u.zk[k] = V.zk[k];
    }
    u.n = f.n_verts;
    rk_update_kernel<<<nblk(u.n), 256, 0, s>>>(u);
    LPM_HIP_CHECK(hipGetLastError());
  }
  u.x = f.face_x;
  u.xes = f.face_x_es;
  u.xcs = f.face_x_cs;
  u.zeta = f.face_zeta;
  u.skip = fskip;
  for (int k = 0; k < 4; ++k) {
    u.k[k] = F.k[k];
    u.zk[k] = F.zk[k];
  }
  // reference face update passes (k1, k2, k4, k4) for the vorticity
  if (flags & LPM_RK4_REFERENCE_COMPAT) u.zk[2] = F.zk[3];
  u.n = f.n_faces;
  rk_update_kernel<<<nblk(u.n), 256, 0, s>>>(u);
  LPM_HIP_CHECK(hipGetLastError());

  eval(map, vert_crd(f), face_crd(f), f.face_zeta);
}

std::string graph_key(const lpm_bve_fields& f, double dt, double Omega,
                      int precision, unsigned flags, hipStream_t s,
                      uint64_t plan_serial) {
  std::ostringstream k;
  k << "bve|" << f.n_verts << '|' << f.vert_x << '|' << f.vert_x_es << '|'
    << f.vert_x_cs << '|' << f.vert_zeta << '|' << f.vert_u << '|'
    << f.vert_u_es << '|' << f.vert_u_cs << '|' << f.n_faces << '|'
    << f.face_x << '|' << f.face_x_es << '|' << f.face_x_cs << '|'
    << f.face_zeta << '|' << f.face_u << '|' << f.face_u_es << '|'
    << f.face_u_cs << '|' << f.face_area << '|'
    << static_cast<const void*>(f.face_mask) << '|';
  k.precision(17);
  k << dt << '|' << Omega << '|' << precision << '|' << flags << '|' << s
    << '|' << plan_serial;
  return k.str();
}

// One RK4 step, directly or through a captured graph.  prepare() allocates
// every workspace the step uses and returns a serial that changes whenever
// the launches it would record change (0 if nothing beyond the fields).
void run_step(lpm_ctx* ctx, const lpm_bve_fields& f, double dt, double Omega,
              int precision, unsigned flags, hipStream_t s,
              const std::function<uint64_t()>& prepare, const EvalFn& eval) {
  StageBufs V = carve(ctx->rk_v, f.n_verts);
  StageBufs F = carve(ctx->rk_f, f.n_faces);
  run_maybe_graph(
      ctx, (flags & LPM_RK4_USE_GRAPH) != 0, s,
      [&]() { return graph_key(f, dt, Omega, precision, flags, s, prepare()); },
      [&]() { issue_step(ctx, f, dt, Omega, flags, V, F, eval, s); });
}

void check_outputs(lpm_ctx* ctx, const lpm_bve_fields& f, unsigned flags,
                   hipStream_t s) {
  if (!(flags & LPM_FLAG_CHECK_FINITE)) return;
  check_finite(ctx, f.vert_u, f.n_verts, f.vert_u_es, f.vert_u_cs, 3, s);
  check_finite(ctx, f.face_u, f.n_faces, f.face_u_es, f.face_u_cs, 3, s);
}

int fail(lpm_ctx* ctx, const std::exception& e) {
  set_error(ctx, e.what());
  const Error* le = dynamic_cast<const Error*>(&e);
  return le ? le->code : LPM_ERR_INTERNAL;
}

// Times every variant of `family` on eval() and records the fastest under
// the tuning key of n_targets.
void autotune(lpm_ctx* ctx, int family, int64_t n_targets, int repeats,
              const std::function<void()>& eval, hipStream_t s, int* best_variant,
              double* best_ms, double* all_ms) {
  LPM_ARG_CHECK(repeats >= 1 && repeats <= 100, "repeats must be in [1, 100]");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const int saved = ctx->forced[family];
  const bool timing = ctx->timing;
  ctx->timing = false;
  auto restore = [&]() {
    ctx->forced[family] = saved;
    ctx->timing = timing;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  try {
    LPM_HIP_CHECK(hipEventCreate(&e0));
    LPM_HIP_CHECK(hipEventCreate(&e1));
    int best = -1;
    float best_t = 0;
    for (int v = 0; v < variant_count(family); ++v) {
      ctx->forced[family] = v;
      eval();  // warm-up (and plan/workspace set-up for this variant)
      LPM_HIP_CHECK(hipEventRecord(e0, s));
      for (int r = 0; r < repeats; ++r) eval();
      LPM_HIP_CHECK(hipEventRecord(e1, s));
      LPM_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0;
      LPM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      ms /= repeats;
      if (all_ms) all_ms[v] = ms;
      if (best < 0 || ms < best_t) {
        best = v;
        best_t = ms;
      }
    }
    restore();
    ctx->tuned[tune_key(ctx, family, n_targets)] = best;
    ++ctx->choice_serial;
    if (best_variant) *best_variant = best;
    if (best_ms) *best_ms = best_t;
  } catch (...) {
    restore();
    throw;
  }
}

}  // namespace

void validate_bve_fields(const lpm_bve_fields* f) {
  LPM_ARG_CHECK(f, "null fields");
  LPM_ARG_CHECK(f->n_verts >= 0 && f->n_faces >= 0, "negative sizes");
  if (f->n_verts > 0) {
    LPM_ARG_CHECK(f->vert_x && f->vert_zeta && f->vert_u,
                  "vertex fields are NULL");
    LPM_ARG_CHECK(f->vert_x_es > 0 && f->vert_x_cs > 0 && f->vert_u_es > 0 &&
                      f->vert_u_cs > 0,
                  "vertex strides must be > 0");
  }
  LPM_ARG_CHECK(f->n_faces > 0, "at least one face is required");
  LPM_ARG_CHECK(f->face_x && f->face_zeta && f->face_u && f->face_area,
                "face fields are NULL");
  LPM_ARG_CHECK(f->face_x_es > 0 && f->face_x_cs > 0 && f->face_u_es > 0 &&
                    f->face_u_cs > 0,
                "face strides must be > 0");
}

void run_maybe_graph(lpm_ctx* ctx, bool use_graph, hipStream_t s,
                     const std::function<std::string()>& key_fn,
                     const std::function<void()>& issue) {
  if (!use_graph) {
    issue();
    return;
  }
  LPM_ARG_CHECK(s != nullptr, "LPM_RK4_USE_GRAPH needs a non-default stream");
  // the key's prepare step allocates; nothing may allocate while capturing
  const std::string key = key_fn() + '|' + std::to_string(ctx->choice_serial);
  auto it = ctx->graphs.find(key);
  if (it != ctx->graphs.end() &&
      it->second->epoch != g_workspace_epoch.load()) {
    // stale (a workspace it records was reallocated): retire it once its
    // last replay has finished
    LPM_HIP_CHECK(hipStreamSynchronize(it->second->stream));
    ctx->graphs.erase(it);
    it = ctx->graphs.end();
  }
  if (it == ctx->graphs.end()) {
    // bounded cache: retire the least recently replayed graph
    while (ctx->graphs.size() >= static_cast<size_t>(LPM_GRAPH_CACHE_MAX)) {
      auto old = ctx->graphs.begin();
      for (auto g = ctx->graphs.begin(); g != ctx->graphs.end(); ++g)
        if (g->second->last_use < old->second->last_use) old = g;
      LPM_HIP_CHECK(hipStreamSynchronize(old->second->stream));
      ctx->graphs.erase(old);
    }
    auto entry = std::make_shared<lpm_ctx::GraphEntry>();
    const bool timing = ctx->timing;
    ctx->timing = false;
    LPM_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      issue();
    } catch (...) {
      ctx->timing = timing;
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    ctx->timing = timing;
    LPM_HIP_CHECK(hipStreamEndCapture(s, &entry->graph));
    LPM_HIP_CHECK(
        hipGraphInstantiate(&entry->exec, entry->graph, nullptr, nullptr, 0));
    entry->epoch = g_workspace_epoch.load();
    it = ctx->graphs.emplace(key, entry).first;
  }
  // a replay is stream-ordered: the host does not wait for it
  it->second->last_use = ++ctx->graph_clock;
  it->second->stream = s;
  LPM_HIP_CHECK(hipGraphLaunch(it->second->exec, s));
}

}  // namespace lpm_amd

using namespace lpm_amd;

namespace {

void check_precision(int precision) {
  LPM_ARG_CHECK(precision == LPM_FP64 || precision == LPM_MIXED,
                "unknown precision");
}

EvalFn single_eval(lpm_ctx* ctx, const lpm_bve_fields& f, int precision,
                   unsigned flags, hipStream_t s) {
  return [ctx, &f, precision, flags, s](const RankMap& map, CrdIn vx, CrdIn fx,
                                        const double* fz) {
    velocity_eval(ctx, f, vx, fx, fz, map, precision, flags, s);
  };
}

EvalFn ensemble_eval(lpm_ctx* ctx, const EnsembleDesc& d,
                     const lpm_bve_fields& f, int precision, unsigned flags,
                     hipStream_t s) {
  return [ctx, &d, &f, precision, flags, s](const RankMap& map, CrdIn vx,
                                            CrdIn fx, const double* fz) {
    ensemble_velocity(ctx, d, f, vx, fx, fz, map, precision, flags, s);
  };
}

EnsembleDesc validate_ensemble(const lpm_bve_ensemble* e) {
  LPM_ARG_CHECK(e, "null ensemble");
  validate_bve_fields(&e->fields);
  return validate_members(e->n_members, e->vert_offsets, e->face_offsets,
                          e->fields.n_verts, e->fields.n_faces);
}

}  // namespace

extern "C" {

int lpm_bve_init_velocity(lpm_ctx* ctx, const lpm_bve_fields* f,
                          int precision, unsigned flags, void* stream) {
  try {
    LPM_ARG_CHECK(ctx, "null context");
    validate_bve_fields(f);
    check_precision(precision);
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const RankMap map = build_rank(ctx, f->face_mask, f->n_faces, s);
    single_eval(ctx, *f, precision, flags, s)(map, vert_crd(*f), face_crd(*f),
                                              f->face_zeta);
    check_outputs(ctx, *f, flags, s);
    return LPM_OK;
  } catch (const std::exception& e) {
    return fail(ctx, e);
  }
}

int lpm_bve_autotune(lpm_ctx* ctx, const lpm_bve_fields* f, int precision,
                     unsigned flags, int repeats, int* best_variant,
                     double* best_ms, double* all_ms, void* stream) {
  try {
    LPM_ARG_CHECK(ctx, "null context");
    validate_bve_fields(f);
    check_precision(precision);
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const int family =
        precision == LPM_MIXED ? KID_BVE_VEL_MIXED : KID_BVE_VEL_F64;
    const RankMap map = build_rank(ctx, f->face_mask, f->n_faces, s);
    const EvalFn ev = single_eval(ctx, *f, precision, flags, s);
    autotune(ctx, family, f->n_verts + f->n_faces, repeats,
             [&]() { ev(map, vert_crd(*f), face_crd(*f), f->face_zeta); }, s,
             best_variant, best_ms, all_ms);
    return LPM_OK;
  } catch (const std::exception& e) {
    return fail(ctx, e);
  }
}

int lpm_bve_rk4_step(lpm_ctx* ctx, const lpm_bve_fields* f, double dt,
                     double Omega, int precision, unsigned flags,
                     void* stream) {
  try {
    LPM_ARG_CHECK(ctx, "null context");
    validate_bve_fields(f);
    check_precision(precision);
    LPM_ARG_CHECK(std::isfinite(dt) && std::isfinite(Omega),
                  "dt and Omega must be finite");
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    auto prepare = [&]() -> uint64_t {
      const RankMap map = build_rank(ctx, f->face_mask, f->n_faces, s);
      pack_sources(ctx, face_crd(*f), f->face_zeta, f->face_area, map,
                   precision, s);
      if (precision == LPM_MIXED)
        target_order(ctx, vert_crd(*f), f->n_verts, nullptr, 1, s);
      const int64_t nt = f->n_verts + f->n_faces;
      const size_t part =
          bve_split_bytes(precision, bve_velocity_variant(ctx, precision, nt), nt);
      if (part) split_workspace(ctx, part, s);  // nullptr: runs unsplit
      return 0;
    };
    run_step(ctx, *f, dt, Omega, precision, flags, s, prepare,
             single_eval(ctx, *f, precision, flags, s));
    check_outputs(ctx, *f, flags, s);
    return LPM_OK;
  } catch (const std::exception& e) {
    return fail(ctx, e);
  }
}

int lpm_bve_ensemble_init_velocity(lpm_ctx* ctx, const lpm_bve_ensemble* e,
                                   int precision, unsigned flags,
                                   void* stream) {
  try {
    LPM_ARG_CHECK(ctx, "null context");
    const EnsembleDesc d = validate_ensemble(e);
    check_precision(precision);
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const lpm_bve_fields& f = e->fields;
    const RankMap map = build_rank(ctx, f.face_mask, f.n_faces, s);
    ensemble_eval(ctx, d, f, precision, flags, s)(map, vert_crd(f), face_crd(f),
                                               f.face_zeta);
    check_outputs(ctx, f, flags, s);
    return LPM_OK;
  } catch (const std::exception& ex) {
    return fail(ctx, ex);
  }
}

int lpm_bve_ensemble_rk4_step(lpm_ctx* ctx, const lpm_bve_ensemble* e,
                              double dt, double Omega, int precision,
                              unsigned flags, void* stream) {
  try {
    LPM_ARG_CHECK(ctx, "null context");
    const EnsembleDesc d = validate_ensemble(e);
    check_precision(precision);
    LPM_ARG_CHECK(std::isfinite(dt) && std::isfinite(Omega),
                  "dt and Omega must be finite");
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const lpm_bve_fields& f = e->fields;
    auto prepare = [&]() -> uint64_t {
      build_rank(ctx, f.face_mask, f.n_faces, s);
      return ensemble_prepare(ctx, d, f, precision, s);
    };
    run_step(ctx, f, dt, Omega, precision, flags, s, prepare,
             ensemble_eval(ctx, d, f, precision, flags, s));
    check_outputs(ctx, f, flags, s);
    return LPM_OK;
  } catch (const std::exception& ex) {
    return fail(ctx, ex);
  }
}

int lpm_bve_ensemble_autotune(lpm_ctx* ctx, const lpm_bve_ensemble* e,
                              int precision, unsigned flags, int repeats,
                              int* best_variant, double* best_ms,
                              double* all_ms, void* stream) {
  try {
    LPM_ARG_CHECK(ctx, "null context");
    const EnsembleDesc d = validate_ensemble(e);
    check_precision(precision);
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const lpm_bve_fields& f = e->fields;
    const int family =
        precision == LPM_MIXED ? KID_BVE_ENS_MIXED : KID_BVE_ENS_F64;
    const RankMap map = build_rank(ctx, f.face_mask, f.n_faces, s);
    const EvalFn ev = ensemble_eval(ctx, d, f, precision, flags, s);
    autotune(ctx, family, f.n_verts + f.n_faces, repeats,
             [&]() { ev(map, vert_crd(f), face_crd(f), f.face_zeta); }, s,
             best_variant, best_ms, all_ms);
    return LPM_OK;
  } catch (const std::exception& ex) {
    return fail(ctx, ex);
  }
}

}  // extern "C"
