// Sphere SWE entry points: direct sums and the RK4 and RK2 steppers.  RK4
// follows the stage layout of SWERK4::advance_timestep
// (src/lpm_swe_rk4_impl.hpp:204-445), RK2 (Heun) that of
// SWERK2::advance_timestep_impl (src/lpm_swe_rk2_impl.hpp:80-257) with the
// surface Laplacian taken from the PSE sums in place of the reference's GMLS
// evaluation; both use the sphere sums and the sphere
// branch of the tendency functors (src/lpm_swe_kernels.hpp:941-1069), zero
// bottom topography (SetSurfaceFromDepth / SetDepthAndSurfaceFromMassAndArea,
// :1079-1140).  Per stage, one fused kernel per particle kind evaluates the
// tendencies, stores the stage increments and forms the next stage's work
// state with its surface.
#include <hip/hip_runtime.h>

#include <cmath>
#include <functional>
#include <sstream>

#include "internal.hpp"

namespace lpm_amd {

namespace {

// Stage storage of one particle kind (SoA, leading dimension n):
//   k[s] (3n), zk[s], sk[s], hk[s] (n each) for s = 0..3;
//   xw (3n), zw, sw, hw (n each).
struct SweStage {
  double* k[4];
  double* zk[4];
  double* sk[4];
  double* hk[4];
  double* xw;
  double* zw;
  double* sw;
  double* hw;
  int64_t n;
};

SweStage carve(DevBuf& buf, int64_t n) {
  const int64_t nn = std::max<int64_t>(n, 1);
  double* p = static_cast<double*>(buf.ensure(sizeof(double) * 30 * nn));
  SweStage b;
  for (int s = 0; s < 4; ++s) {
    b.k[s] = p + (6 * s) * nn;
    b.zk[s] = p + (6 * s + 3) * nn;
    b.sk[s] = p + (6 * s + 4) * nn;
    b.hk[s] = p + (6 * s + 5) * nn;
  }
  b.xw = p + 24 * nn;
  b.zw = p + 27 * nn;
  b.sw = p + 28 * nn;
  b.hw = p + 29 * nn;
  b.n = n;
  return b;
}

struct SweStageArgs {
  // stage input (state for stage 0, work arrays afterwards)
  const double* x;
  int64_t xes, xcs;
  const double* zeta;
  const double* sigma;
  const double* ha;  // depth (vertices) or area (faces)
  // sums at the stage input
  const double* u;
  int64_t ues, ucs;
  const double* ddot;
  const double* laps;
  // state (for the next work state)
  const double* x0;
  int64_t x0es, x0cs;
  const double* zeta0;
  const double* sigma0;
  const double* ha0;
  // outputs
  double* k;
  double* zk;
  double* sk;
  double* hk;
  double* xw;  // nullptr after the last stage
  double* zw;
  double* sw;
  double* hw;
  double* surf;       // surface of the next work state
  double* face_h;     // faces: depth = mass / area of the next work state
  const double* mass;
  const uint8_t* skip;
  const uint8_t* mask;  // faces: divided faces keep h, s
  int64_t n;
  double dt, Omega, g, a;
  int is_face;
  int store_k;  // 0: leave k[s] (position increment) at zero
  int renorm;   // project the work positions onto the unit sphere
};

// x / |x| in FP64 (LPM_RK4_RENORMALIZE).
__device__ __forceinline__ void project_unit(double& w0, double& w1,
                                             double& w2) {
  const double r = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2);
  w0 *= r;
  w1 *= r;
  w2 *= r;
}

// Tendencies (sphere branch), stage increments and the next work state.
__global__ __launch_bounds__(256) void swe_stage_kernel(SweStageArgs A) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  if (A.skip && A.skip[i]) return;
  const int64_t n = A.n;
  const double* xi = A.x + i * A.xes;
  const double x0 = xi[0], x1 = xi[A.xcs], x2 = xi[2 * A.xcs];
  const double* ui = A.u + i * A.ues;
  const double u0 = ui[0], u1 = ui[A.ucs], u2 = ui[2 * A.ucs];
  const double zeta = A.zeta[i], sigma = A.sigma[i], ha = A.ha[i];
  const double f = 2.0 * A.Omega * x2;
  const double dfdt = 2.0 * A.Omega * u2;
  const double gfxu = -2.0 * A.Omega * (-u0 * x1 + u1 * x0);
  const double dz = (-dfdt - (zeta + f) * sigma) * A.dt;
  const double ds = (f * zeta + gfxu - A.ddot[i] - A.g * A.laps[i] -
                     (u0 * u0 + u1 * u1 + u2 * u2)) * A.dt;
  const double dh = A.is_face ? (sigma * ha) * A.dt : (-sigma * ha) * A.dt;
  const double k0 = A.store_k ? A.dt * u0 : 0.0;
  const double k1 = A.store_k ? A.dt * u1 : 0.0;
  const double k2 = A.store_k ? A.dt * u2 : 0.0;
  A.k[i] = k0;
  A.k[i + n] = k1;
  A.k[i + 2 * n] = k2;
  A.zk[i] = dz;
  A.sk[i] = ds;
  A.hk[i] = dh;
  if (A.xw) {
    const double* y = A.x0 + i * A.x0es;
    double w0 = y[0] + A.a * k0;
    double w1 = y[A.x0cs] + A.a * k1;
    double w2 = y[2 * A.x0cs] + A.a * k2;
    if (A.renorm) project_unit(w0, w1, w2);
    A.xw[i] = w0;
    A.xw[i + n] = w1;
    A.xw[i + 2 * n] = w2;
    A.zw[i] = A.zeta0[i] + A.a * dz;
    A.sw[i] = A.sigma0[i] + A.a * ds;
    const double hw = A.ha0[i] + A.a * dh;
    A.hw[i] = hw;
    if (!A.is_face) {
      A.surf[i] = hw;
    } else if (!(A.mask && A.mask[i])) {
      const double h = A.mass[i] / hw;
      A.face_h[i] = h;
      A.surf[i] = h;
    }
  }
}

struct SweUpdateArgs {
  double* x;
  int64_t xes, xcs;
  double* zeta;
  double* sigma;
  double* ha;
  const double* k[4];
  const double* zk[4];
  const double* sk[4];
  const double* hk[4];
  double* surf;
  double* face_h;
  const double* mass;
  const uint8_t* skip;
  const uint8_t* mask;
  int64_t n;
  int is_face;
  int order;  // 4: SWERK4Update weights; 2: Heun, x += (x1/2 + x2/2)
  int renorm; // project the new positions onto the unit sphere
};

// SWERK4Update (or the RK2 update) and the surface of the new state.
__global__ __launch_bounds__(256) void swe_update_kernel(SweUpdateArgs A) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  if (A.skip && A.skip[i]) return;
  constexpr double sixth = 1.0 / 6.0, third = 1.0 / 3.0;
  const int64_t n = A.n;
  double* x = A.x + i * A.xes;
  double v[3];
  if (A.order == 2) {
    // KokkosBlas::update(0.5, k1, 0.5, k2, 1, state) of the reference's RK2
    for (int c = 0; c < 3; ++c) {
      const int64_t o = i + c * n;
      v[c] = x[c * A.xcs] + (0.5 * A.k[0][o] + 0.5 * A.k[1][o]);
    }
    if (A.renorm) project_unit(v[0], v[1], v[2]);
    for (int c = 0; c < 3; ++c) x[c * A.xcs] = v[c];
    A.zeta[i] += 0.5 * A.zk[0][i] + 0.5 * A.zk[1][i];
    A.sigma[i] += 0.5 * A.sk[0][i] + 0.5 * A.sk[1][i];
    const double h = A.ha[i] + (0.5 * A.hk[0][i] + 0.5 * A.hk[1][i]);
    A.ha[i] = h;
    if (!A.is_face) {
      A.surf[i] = h;
    } else if (!(A.mask && A.mask[i])) {
      const double d = A.mass[i] / h;
      A.face_h[i] = d;
      A.surf[i] = d;
    }
    return;
  }
  for (int c = 0; c < 3; ++c) {
    const int64_t o = i + c * n;
    v[c] = x[c * A.xcs] + (sixth * (A.k[0][o] + A.k[3][o]) +
                           third * (A.k[1][o] + A.k[2][o]));
  }
  if (A.renorm) project_unit(v[0], v[1], v[2]);
  for (int c = 0; c < 3; ++c) x[c * A.xcs] = v[c];
  A.zeta[i] += sixth * (A.zk[0][i] + A.zk[3][i]) + third * (A.zk[1][i] + A.zk[2][i]);
  A.sigma[i] += sixth * (A.sk[0][i] + A.sk[3][i]) + third * (A.sk[1][i] + A.sk[2][i]);
  const double h = A.ha[i] + (sixth * (A.hk[0][i] + A.hk[3][i]) +
                              third * (A.hk[1][i] + A.hk[2][i]));
  A.ha[i] = h;
  if (!A.is_face) {
    A.surf[i] = h;
  } else if (!(A.mask && A.mask[i])) {
    const double d = A.mass[i] / h;
    A.face_h[i] = d;
    A.surf[i] = d;
  }
}

inline int nblk(int64_t n) { return static_cast<int>((n + 255) / 256); }

void validate(const lpm_swe_fields* f) {
  LPM_ARG_CHECK(f, "null fields");
  LPM_ARG_CHECK(f->n_verts >= 0 && f->n_faces > 0, "bad sizes");
  if (f->n_verts > 0)
    LPM_ARG_CHECK(f->vert_x && f->vert_zeta && f->vert_sigma && f->vert_h &&
                      f->vert_u && f->vert_ddot && f->vert_laps && f->vert_surf &&
                      f->vert_x_es > 0 && f->vert_x_cs > 0 && f->vert_u_es > 0 &&
                      f->vert_u_cs > 0,
                  "vertex fields incomplete");
  LPM_ARG_CHECK(f->face_x && f->face_zeta && f->face_sigma && f->face_area &&
                    f->face_mass && f->face_h && f->face_u && f->face_ddot &&
                    f->face_laps && f->face_surf && f->face_x_es > 0 &&
                    f->face_x_cs > 0 && f->face_u_es > 0 && f->face_u_cs > 0,
                "face fields incomplete");
}

// Sums at positions (vx, fx) with face sources (fx, fzeta, fsigma, farea,
// surf) into the fields' u, ddot, laps.
using SweEvalFn = std::function<void(const RankMap&, CrdIn vx, CrdIn fx,
                                     const double* fzeta, const double* fsigma,
                                     const double* farea)>;

void sums_eval(lpm_ctx* ctx, const lpm_swe_fields& f, CrdIn vx, CrdIn fx,
               const double* fzeta, const double* fsigma, const double* farea,
               const RankMap& map, double eps, double pse_eps, int precision,
               unsigned flags, hipStream_t s) {
  SwePacked ps = swe_pack_sources(ctx, fx, fzeta, fsigma, farea, f.face_surf,
                                  map, s);
  SweTargets segs[2];
  segs[0].base.x = vx;
  segs[0].base.n = f.n_verts;
  // vertex slots in the spatial order of the state positions (the stage
  // positions of a step are close to them)
  if (precision == LPM_MIXED)
    segs[0].base.index =
        target_order(ctx, CrdIn{f.vert_x, f.vert_x_es, f.vert_x_cs},
                     f.n_verts, nullptr, 1, s);
  segs[0].base.out = CrdOut{f.vert_u, f.vert_u_es, f.vert_u_cs};
  segs[0].surf = f.vert_surf;
  segs[0].ddot = f.vert_ddot;
  segs[0].laps = f.vert_laps;
  segs[1].base = face_targets(fx, map, f.face_mask,
                              (flags & LPM_FLAG_SKIP_MASKED_TARGETS) != 0,
                              CrdOut{f.face_u, f.face_u_es, f.face_u_cs},
                              nullptr);
  segs[1].surf = f.face_surf;
  segs[1].ddot = f.face_ddot;
  segs[1].laps = f.face_laps;
  launch_swe_sums(ctx, ps, segs, 2, eps, pse_eps, precision, s);
}

CrdIn swe_vert_crd(const lpm_swe_fields& f) {
  return CrdIn{f.vert_x, f.vert_x_es, f.vert_x_cs};
}
CrdIn swe_face_crd(const lpm_swe_fields& f) {
  return CrdIn{f.face_x, f.face_x_es, f.face_x_cs};
}

// One step of order 4 (RK4) or 2 (Heun RK2).  RK2: the work state is
// state + stage-1 increments (KokkosBlas::update(1, x, 1, x1, 0, xwork)),
// its sums and tendencies give the stage-2 increments.
void issue_swe_step(lpm_ctx* ctx, const lpm_swe_fields& f, double dt,
                    double Omega, double g, unsigned flags, int order,
                    SweStage& V, SweStage& F, const SweEvalFn& eval,
                    hipStream_t s) {
  const uint8_t* fskip =
      (flags & LPM_FLAG_SKIP_MASKED_TARGETS) ? f.face_mask : nullptr;
  const bool compat = order == 4 && (flags & LPM_RK4_REFERENCE_COMPAT) != 0;
  const RankMap map = build_rank(ctx, f.face_mask, f.n_faces, s);
  const double coef[3] = {0.5, 0.5, 1.0};
  const int nstage = order == 2 ? 2 : 4;
  for (int stage = 0; stage < nstage; ++stage) {
    if (stage > 0)
      eval(map, CrdIn{V.xw, 1, V.n}, CrdIn{F.xw, 1, F.n}, F.zw, F.sw, F.hw);
    for (int kind = 0; kind < 2; ++kind) {
      const bool face = kind == 1;
      SweStage& B = face ? F : V;
      const int64_t n = face ? f.n_faces : f.n_verts;
      if (n == 0) continue;
      SweStageArgs a{};
      if (stage == 0) {
        a.x = face ? f.face_x : f.vert_x;
        a.xes = face ? f.face_x_es : f.vert_x_es;
        a.xcs = face ? f.face_x_cs : f.vert_x_cs;
        a.zeta = face ? f.face_zeta : f.vert_zeta;
        a.sigma = face ? f.face_sigma : f.vert_sigma;
        a.ha = face ? f.face_area : f.vert_h;
      } else {
        a.x = B.xw;
        a.xes = 1;
        a.xcs = B.n;
        a.zeta = B.zw;
        a.sigma = B.sw;
        a.ha = B.hw;
      }
      a.u = face ? f.face_u : f.vert_u;
      a.ues = face ? f.face_u_es : f.vert_u_es;
      a.ucs = face ? f.face_u_cs : f.vert_u_cs;
      a.ddot = face ? f.face_ddot : f.vert_ddot;
      a.laps = face ? f.face_laps : f.vert_laps;
      a.x0 = face ? f.face_x : f.vert_x;
      a.x0es = face ? f.face_x_es : f.vert_x_es;
      a.x0cs = face ? f.face_x_cs : f.vert_x_cs;
      a.zeta0 = face ? f.face_zeta : f.vert_zeta;
      a.sigma0 = face ? f.face_sigma : f.vert_sigma;
      a.ha0 = face ? f.face_area : f.vert_h;
      a.k = B.k[stage];
      a.zk = B.zk[stage];
      a.sk = B.sk[stage];
      a.hk = B.hk[stage];
      if (stage < nstage - 1) {
        a.xw = B.xw;
        a.zw = B.zw;
        a.sw = B.sw;
        a.hw = B.hw;
        a.a = order == 2 ? 1.0 : coef[stage];
      }
      a.surf = face ? f.face_surf : f.vert_surf;
      a.face_h = f.face_h;
      a.mass = f.face_mass;
      a.skip = face ? fskip : nullptr;
      a.mask = face ? f.face_mask : nullptr;
      a.n = n;
      a.dt = dt;
      a.Omega = Omega;
      a.g = g;
      a.is_face = face ? 1 : 0;
      // the reference never assigns the stage-4 position increment
      a.store_k = (compat && stage == 3) ? 0 : 1;
      a.renorm = (flags & LPM_RK4_RENORMALIZE) ? 1 : 0;
      swe_stage_kernel<<<nblk(n), 256, 0, s>>>(a);
      LPM_HIP_CHECK(hipGetLastError());
    }
  }
  for (int kind = 0; kind < 2; ++kind) {
    const bool face = kind == 1;
    SweStage& B = face ? F : V;
    const int64_t n = face ? f.n_faces : f.n_verts;
    if (n == 0) continue;
    SweUpdateArgs u{};
    u.x = face ? f.face_x : f.vert_x;
    u.xes = face ? f.face_x_es : f.vert_x_es;
    u.xcs = face ? f.face_x_cs : f.vert_x_cs;
    u.zeta = face ? f.face_zeta : f.vert_zeta;
    u.sigma = face ? f.face_sigma : f.vert_sigma;
    u.ha = face ? f.face_area : f.vert_h;
    for (int k = 0; k < 4; ++k) {
      u.k[k] = B.k[k];
      u.zk[k] = B.zk[k];
      u.sk[k] = B.sk[k];
      u.hk[k] = B.hk[k];
    }
    u.surf = face ? f.face_surf : f.vert_surf;
    u.face_h = f.face_h;
    u.mass = f.face_mass;
    u.skip = face ? fskip : nullptr;
    u.mask = face ? f.face_mask : nullptr;
    u.n = n;
    u.is_face = face ? 1 : 0;
    u.order = order;
    u.renorm = (flags & LPM_RK4_RENORMALIZE) ? 1 : 0;
    swe_update_kernel<<<nblk(n), 256, 0, s>>>(u);
    LPM_HIP_CHECK(hipGetLastError());
  }
  eval(map, swe_vert_crd(f), swe_face_crd(f), f.face_zeta, f.face_sigma,
       f.face_area);
}

std::string swe_graph_key(const lpm_swe_fields& f, double dt, double Omega,
                          double g, double eps, double pse_eps, int precision,
                          unsigned flags, int order, hipStream_t s,
                          uint64_t serial) {
  std::ostringstream k;
  k << "swe" << order << '|' << f.n_verts << '|' << f.vert_x << '|' << f.vert_x_es << '|'
    << f.vert_x_cs << '|' << f.vert_zeta << '|' << f.vert_sigma << '|'
    << f.vert_h << '|' << f.vert_u << '|' << f.vert_u_es << '|'
    << f.vert_u_cs << '|' << f.vert_ddot << '|' << f.vert_laps << '|'
    << f.vert_surf << '|' << f.n_faces << '|' << f.face_x << '|'
    << f.face_x_es << '|' << f.face_x_cs << '|' << f.face_zeta << '|'
    << f.face_sigma << '|' << f.face_area << '|' << f.face_mass << '|'
    << f.face_h << '|' << f.face_u << '|' << f.face_u_es << '|'
    << f.face_u_cs << '|' << f.face_ddot << '|' << f.face_laps << '|'
    << f.face_surf << '|' << static_cast<const void*>(f.face_mask) << '|';
  k.precision(17);
  k << dt << '|' << Omega << '|' << g << '|' << eps << '|' << pse_eps << '|'
    << precision << '|' << flags << '|' << s << '|' << serial;
  return k.str();
}

void check_step_args(lpm_ctx* ctx, int precision, double dt, double Omega,
                     double g, double eps, double pse_eps) {
  LPM_ARG_CHECK(ctx, "null context");
  LPM_ARG_CHECK(precision == LPM_FP64 || precision == LPM_MIXED,
                "unknown precision");
  LPM_ARG_CHECK(std::isfinite(dt) && std::isfinite(Omega) && std::isfinite(g),
                "dt, Omega, g must be finite");
  LPM_ARG_CHECK(eps >= 0 && pse_eps >= 0, "eps, pse_eps must be >= 0");
}

void check_swe_outputs(lpm_ctx* ctx, const lpm_swe_fields& f, unsigned flags,
                       hipStream_t s) {
  if (!(flags & LPM_FLAG_CHECK_FINITE)) return;
  check_finite(ctx, f.vert_u, f.n_verts, f.vert_u_es, f.vert_u_cs, 3, s);
  check_finite(ctx, f.face_u, f.n_faces, f.face_u_es, f.face_u_cs, 3, s);
}

SweEvalFn single_swe_eval(lpm_ctx* ctx, const lpm_swe_fields& f, double eps,
                          double pse_eps, int precision, unsigned flags,
                          hipStream_t s) {
  return [=, &f](const RankMap& map, CrdIn vx, CrdIn fx, const double* fz,
                 const double* fsig, const double* fa) {
    sums_eval(ctx, f, vx, fx, fz, fsig, fa, map, eps, pse_eps, precision,
              flags, s);
  };
}

SweEvalFn ensemble_swe_eval(lpm_ctx* ctx, const EnsembleDesc& d,
                            const lpm_swe_fields& f, double eps,
                            double pse_eps, int precision, unsigned flags,
                            hipStream_t s) {
  return [=, &d, &f](const RankMap& map, CrdIn vx, CrdIn fx, const double* fz,
                     const double* fsig, const double* fa) {
    swe_ensemble_sums(ctx, d, f, vx, fx, fz, fsig, fa, map, eps, pse_eps,
                      precision, flags, s);
  };
}

EnsembleDesc validate_swe_ensemble(const lpm_swe_ensemble* e) {
  LPM_ARG_CHECK(e, "null ensemble");
  validate(&e->fields);
  return validate_members(e->n_members, e->vert_offsets, e->face_offsets,
                          e->fields.n_verts, e->fields.n_faces);
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

int lpm_swe_sphere_sums(lpm_ctx* ctx, const double* tgt_x, int64_t tgt_es,
                        int64_t tgt_cs, int64_t n_tgt, const double* tgt_surf,
                        const double* src_x, int64_t src_es, int64_t src_cs,
                        const double* src_zeta, const double* src_sigma,
                        const double* src_area, const uint8_t* src_mask,
                        const double* src_surf, int64_t n_src, int collocated,
                        double eps, double pse_eps, int precision,
                        unsigned flags, double* out_u, int64_t out_es,
                        int64_t out_cs, double* out_ddot, double* out_laps,
                        void* stream) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx, "null context");
    LPM_ARG_CHECK(n_tgt >= 0 && n_src >= 0, "negative size");
    LPM_ARG_CHECK(precision == LPM_FP64 || precision == LPM_MIXED,
                  "unknown precision");
    LPM_ARG_CHECK(eps >= 0 && pse_eps >= 0, "eps, pse_eps must be >= 0");
    if (n_tgt > 0) {
      LPM_ARG_CHECK(tgt_x && tgt_es > 0 && tgt_cs > 0, "bad tgt_x");
      LPM_ARG_CHECK(!out_u || (out_es > 0 && out_cs > 0), "bad out_u strides");
    }
    if (n_src > 0) {
      LPM_ARG_CHECK(src_x && src_es > 0 && src_cs > 0, "bad src_x");
      LPM_ARG_CHECK(src_zeta && src_sigma && src_area,
                    "src_zeta/src_sigma/src_area are NULL");
    }
    if (pse_eps > 0)
      LPM_ARG_CHECK((tgt_surf || n_tgt == 0) && (src_surf || n_src == 0),
                    "PSE needs tgt_surf and src_surf");
    if (collocated)
      LPM_ARG_CHECK(n_tgt == n_src,
                    "collocated: targets must be the sources (n_tgt == n_src)");
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const RankMap map = build_rank(ctx, src_mask, n_src, s);
    SwePacked ps = swe_pack_sources(ctx, CrdIn{src_x, src_es, src_cs},
                                    src_zeta, src_sigma, src_area, src_surf,
                                    map, s);
    SweTargets t;
    if (collocated) {
      t.base = face_targets(CrdIn{tgt_x, tgt_es, tgt_cs}, map, src_mask,
                            (flags & LPM_FLAG_SKIP_MASKED_TARGETS) != 0,
                            CrdOut{out_u, out_es, out_cs}, nullptr);
    } else {
      t.base.x = CrdIn{tgt_x, tgt_es, tgt_cs};
      t.base.n = n_tgt;
      t.base.out = CrdOut{out_u, out_es, out_cs};
      if (precision == LPM_MIXED)
        t.base.index = target_order(ctx, t.base.x, n_tgt, nullptr, 1, s);
    }
    t.surf = tgt_surf;
    t.ddot = out_ddot;
    t.laps = out_laps;
    launch_swe_sums(ctx, ps, &t, 1, eps, pse_eps, precision, s);
    if ((flags & LPM_FLAG_CHECK_FINITE) && out_u)
      check_finite(ctx, out_u, n_tgt, out_es, out_cs, 3, s);
  });
}

int lpm_swe_init_sums(lpm_ctx* ctx, const lpm_swe_fields* f, double eps,
                      double pse_eps, int precision, unsigned flags,
                      void* stream) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx, "null context");
    validate(f);
    LPM_ARG_CHECK(precision == LPM_FP64 || precision == LPM_MIXED,
                  "unknown precision");
    LPM_ARG_CHECK(eps >= 0 && pse_eps >= 0, "eps, pse_eps must be >= 0");
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const RankMap map = build_rank(ctx, f->face_mask, f->n_faces, s);
    sums_eval(ctx, *f, swe_vert_crd(*f), swe_face_crd(*f), f->face_zeta,
              f->face_sigma, f->face_area, map, eps, pse_eps, precision, flags,
              s);
  });
}

}  // extern "C"

namespace lpm_amd {
namespace {

int swe_step(lpm_ctx* ctx, const lpm_swe_fields* f, double dt, double Omega,
             double g, double eps, double pse_eps, int precision,
             unsigned flags, int order, void* stream) {
  return guarded(ctx, [&] {
    check_step_args(ctx, precision, dt, Omega, g, eps, pse_eps);
    validate(f);
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    SweStage V = carve(ctx->swe_v, f->n_verts);
    SweStage F = carve(ctx->swe_f, f->n_faces);
    const SweEvalFn ev =
        single_swe_eval(ctx, *f, eps, pse_eps, precision, flags, s);
    run_maybe_graph(
        ctx, (flags & LPM_RK4_USE_GRAPH) != 0, s,
        [&]() {
          const RankMap map = build_rank(ctx, f->face_mask, f->n_faces, s);
          swe_pack_sources(ctx, swe_face_crd(*f), f->face_zeta, f->face_sigma,
                           f->face_area, f->face_surf, map, s);
          swe_feature_workspace(ctx, f->n_faces, f->n_verts + f->n_faces,
                                precision);
          if (precision == LPM_MIXED)
            target_order(ctx, swe_vert_crd(*f), f->n_verts, nullptr, 1, s);
          return swe_graph_key(*f, dt, Omega, g, eps, pse_eps, precision,
                               flags, order, s, 0);
        },
        [&]() {
          issue_swe_step(ctx, *f, dt, Omega, g, flags, order, V, F, ev, s);
        });
    check_swe_outputs(ctx, *f, flags, s);
  });
}

int swe_ensemble_step(lpm_ctx* ctx, const lpm_swe_ensemble* e, double dt,
                      double Omega, double g, double eps, double pse_eps,
                      int precision, unsigned flags, int order, void* stream) {
  return guarded(ctx, [&] {
    check_step_args(ctx, precision, dt, Omega, g, eps, pse_eps);
    const EnsembleDesc d = validate_swe_ensemble(e);
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const lpm_swe_fields& f = e->fields;
    SweStage V = carve(ctx->swe_v, f.n_verts);
    SweStage F = carve(ctx->swe_f, f.n_faces);
    const SweEvalFn ev =
        ensemble_swe_eval(ctx, d, f, eps, pse_eps, precision, flags, s);
    run_maybe_graph(
        ctx, (flags & LPM_RK4_USE_GRAPH) != 0, s,
        [&]() {
          build_rank(ctx, f.face_mask, f.n_faces, s);
          return swe_graph_key(f, dt, Omega, g, eps, pse_eps, precision, flags,
                               order, s,
                               swe_ensemble_prepare(ctx, d, f, precision, s));
        },
        [&]() {
          issue_swe_step(ctx, f, dt, Omega, g, flags, order, V, F, ev, s);
        });
    check_swe_outputs(ctx, f, flags, s);
  });
}

}  // namespace
}  // namespace lpm_amd

extern "C" {

int lpm_swe_rk4_step(lpm_ctx* ctx, const lpm_swe_fields* f, double dt,
                     double Omega, double g, double eps, double pse_eps,
                     int precision, unsigned flags, void* stream) {
  return swe_step(ctx, f, dt, Omega, g, eps, pse_eps, precision, flags, 4,
                  stream);
}

int lpm_swe_rk2_step(lpm_ctx* ctx, const lpm_swe_fields* f, double dt,
                     double Omega, double g, double eps, double pse_eps,
                     int precision, unsigned flags, void* stream) {
  return swe_step(ctx, f, dt, Omega, g, eps, pse_eps, precision, flags, 2,
                  stream);
}

int lpm_swe_ensemble_init_sums(lpm_ctx* ctx, const lpm_swe_ensemble* e,
                               double eps, double pse_eps, int precision,
                               unsigned flags, void* stream) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx, "null context");
    const EnsembleDesc d = validate_swe_ensemble(e);
    LPM_ARG_CHECK(precision == LPM_FP64 || precision == LPM_MIXED,
                  "unknown precision");
    LPM_ARG_CHECK(eps >= 0 && pse_eps >= 0, "eps, pse_eps must be >= 0");
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const lpm_swe_fields& f = e->fields;
    const RankMap map = build_rank(ctx, f.face_mask, f.n_faces, s);
    swe_ensemble_sums(ctx, d, f, swe_vert_crd(f), swe_face_crd(f),
                      f.face_zeta, f.face_sigma, f.face_area, map, eps,
                      pse_eps, precision, flags, s);
    check_swe_outputs(ctx, f, flags, s);
  });
}

int lpm_swe_ensemble_rk4_step(lpm_ctx* ctx, const lpm_swe_ensemble* e,
                              double dt, double Omega, double g, double eps,
                              double pse_eps, int precision, unsigned flags,
                              void* stream) {
  return swe_ensemble_step(ctx, e, dt, Omega, g, eps, pse_eps, precision,
                           flags, 4, stream);
}

int lpm_swe_ensemble_rk2_step(lpm_ctx* ctx, const lpm_swe_ensemble* e,
                              double dt, double Omega, double g, double eps,
                              double pse_eps, int precision, unsigned flags,
                              void* stream) {
  return swe_ensemble_step(ctx, e, dt, Omega, g, eps, pse_eps, precision,
                           flags, 2, stream);
}

}  // extern "C"
