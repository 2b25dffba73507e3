/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C, FP64 restatement of LPM's spherical direct sums, RK4 steppers and
 * galleries (pbosler/lpm).  It is the checker for the HIP product path: only
 * tests/, the smoke check and the benchmark's CPU-baseline leg may load it.
 * Nothing in the product links or calls it.
 *
 * Semantics are "reference-faithful": every target visits every stored face
 * j = 0..n_src-1 in order, skips masked (divided) faces and, for collocated
 * targets, j == i; each target's sum is a serial left-to-right reduction, the
 * order of a Kokkos team reduction of team size 1 (Kokkos-OpenMP backend).
 * Targets are spread over OpenMP threads.
 *
 * Coordinates are AoS (n,3) row-major, the Kokkos-OpenMP (LayoutRight)
 * layout of View<Real*[3]>.
 *
 * Parity status of each function is stated at its definition; functions with
 * no golden vector, closed form or brute-force check say "parity unpinned".
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_PI 3.1415926535897932384626433832795027975

static inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3(double* c, const double* a, const double* b) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

int ora_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ BVE --
 * biot_savart, src/lpm_sphere_functions.hpp:44-57:
 *   u = (x cross y) * (-zeta*A) / (4 pi (1 - y.x)).
 * Pinned by: solid-body and RH5-4 closed forms of the direct sum, tangency. */
void ora_biot_savart(double* u, const double* x, const double* y, double zeta,
                     double area) {
  cross3(u, x, y);
  const double strength = -zeta * area / (4 * ORA_PI * (1 - dot3(y, x)));
  for (int j = 0; j < 3; ++j) u[j] *= strength;
}

/* greens_fn, src/lpm_sphere_functions.hpp:21-29:
 *   psi = log(1 - x.y) * (-zeta*A) / (4 pi).
 * Pinned by: solid-body stream function, spherical-harmonic eigenrelation. */
double ora_greens_fn(const double* x, const double* y, double zeta,
                     double area) {
  const double circ = -zeta * area;
  return log(1 - dot3(x, y)) * circ / (4 * ORA_PI);
}

/* BVEVertexVelocity / BVEFaceVelocity with VelocityReduceDistinct /
 * VelocityReduceCollocated, src/lpm_bve_sphere_kernels.hpp:53-82, 179-211,
 * 249-274, 365-394.  collocated != 0: targets are the sources themselves
 * (tx == sx, nt == ns) and j == i is skipped.  smask may be NULL (no face
 * divided).  tgt_mask (nullable): targets whose flag is set are skipped and
 * left untouched. */
void ora_bve_velocity(const double* tx, int64_t nt, const double* sx,
                      const double* szeta, const double* sarea,
                      const uint8_t* smask, int64_t ns, int collocated,
                      const uint8_t* tgt_mask, double* u, int nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < nt; ++i) {
    if (tgt_mask && tgt_mask[i]) continue;
    double acc[3] = {0, 0, 0};
    const double* x = tx + 3 * i;
    for (int64_t j = 0; j < ns; ++j) {
      double v[3] = {0, 0, 0};
      if (!(smask && smask[j]) && !(collocated && i == j))
        ora_biot_savart(v, x, sx + 3 * j, szeta[j], sarea[j]);
      acc[0] += v[0];
      acc[1] += v[1];
      acc[2] += v[2];
    }
    u[3 * i + 0] = acc[0];
    u[3 * i + 1] = acc[1];
    u[3 * i + 2] = acc[2];
  }
}

/* BVEVertexStreamFn / BVEFaceStreamFn with StreamReduceDistinct /
 * StreamReduceCollocated, src/lpm_bve_sphere_kernels.hpp:20-48, 141-170,
 * 218-242, 329-356. */
void ora_bve_stream_fn(const double* tx, int64_t nt, const double* sx,
                       const double* szeta, const double* sarea,
                       const uint8_t* smask, int64_t ns, int collocated,
                       const uint8_t* tgt_mask, double* psi, int nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < nt; ++i) {
    if (tgt_mask && tgt_mask[i]) continue;
    double acc = 0;
    for (int64_t j = 0; j < ns; ++j) {
      double p = 0;
      if (!(smask && smask[j]) && !(collocated && i == j))
        p = ora_greens_fn(tx + 3 * i, sx + 3 * j, szeta[j], sarea[j]);
      acc += p;
    }
    psi[i] = acc;
  }
}

/* Extended-precision (long double, Kahan-compensated) BVE velocity: the
 * brute-force check of the FP64 sums on small meshes. */
void ora_bve_velocity_xp(const double* tx, int64_t nt, const double* sx,
                         const double* szeta, const double* sarea,
                         const uint8_t* smask, int64_t ns, int collocated,
                         double* u, int nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < nt; ++i) {
    long double acc[3] = {0, 0, 0}, comp[3] = {0, 0, 0};
    const double* x = tx + 3 * i;
    for (int64_t j = 0; j < ns; ++j) {
      if ((smask && smask[j]) || (collocated && i == j)) continue;
      const double* y = sx + 3 * j;
      const long double c[3] = {
          (long double)x[1] * y[2] - (long double)x[2] * y[1],
          (long double)x[2] * y[0] - (long double)x[0] * y[2],
          (long double)x[0] * y[1] - (long double)x[1] * y[0]};
      const long double t = (long double)x[0] * y[0] +
                            (long double)x[1] * y[1] +
                            (long double)x[2] * y[2];
      const long double s =
          -(long double)szeta[j] * sarea[j] / (4 * 3.14159265358979323846264338327950288L * (1 - t));
      for (int k = 0; k < 3; ++k) {
        const long double yk = c[k] * s - comp[k];
        const long double tk = acc[k] + yk;
        comp[k] = (tk - acc[k]) - yk;
        acc[k] = tk;
      }
    }
    for (int k = 0; k < 3; ++k) u[3 * i + k] = (double)acc[k];
  }
}

/* BVEVorticityTendency, src/lpm_bve_sphere_kernels.hpp:399-413:
 *   dzeta = -2 Omega u_z dt. */
static inline double bve_tendency(const double* u, double dt, double Omega) {
  return -2.0 * Omega * u[2] * dt;
}

/* One BVERK4::advance_timestep, src/lpm_bve_rk4_impl.hpp:63-167, with the
 * stage buffers of src/lpm_bve_rk4.cpp:5-39.  On entry vu/fu hold the velocity
 * at the current positions; on exit the velocity at the new positions.
 * compat != 0 reproduces the reference's face-vorticity update, which passes
 * (k1, k2, k4, k4) (src/lpm_bve_rk4_impl.hpp:152-157); compat == 0 uses the
 * classical (k1, k2, k3, k4).  renorm != 0 projects every stage position and
 * the updated position back onto the unit sphere (not done by the reference).
 * Pinned by: solid-body rotation positions, RK4 self-convergence order. */
void ora_bve_rk4_step(double* vx, double* vzeta, double* vu, int64_t nv,
                      double* fx, double* fzeta, double* fu,
                      const double* farea, const uint8_t* fmask, int64_t nf,
                      double dt, double Omega, int compat, int renorm,
                      int nthreads) {
  const double sixth = 1.0 / 6.0, third = 1.0 / 3.0;
  double* vk[4];
  double* vzk[4];
  double* fk[4];
  double* fzk[4];
  for (int s = 0; s < 4; ++s) {
    vk[s] = (double*)malloc(sizeof(double) * 3 * (nv > 0 ? nv : 1));
    vzk[s] = (double*)malloc(sizeof(double) * (nv > 0 ? nv : 1));
    fk[s] = (double*)malloc(sizeof(double) * 3 * (nf > 0 ? nf : 1));
    fzk[s] = (double*)malloc(sizeof(double) * (nf > 0 ? nf : 1));
  }
  double* vxw = (double*)malloc(sizeof(double) * 3 * (nv > 0 ? nv : 1));
  double* vzw = (double*)malloc(sizeof(double) * (nv > 0 ? nv : 1));
  double* fxw = (double*)malloc(sizeof(double) * 3 * (nf > 0 ? nf : 1));
  double* fzw = (double*)malloc(sizeof(double) * (nf > 0 ? nf : 1));
  const double coef[4] = {0.5, 0.5, 1.0, 0.0};

  for (int stage = 0; stage < 4; ++stage) {
    if (stage > 0) {
      ora_bve_velocity(vxw, nv, fxw, fzw, farea, fmask, nf, 0, NULL, vu,
                       nthreads);
      ora_bve_velocity(fxw, nf, fxw, fzw, farea, fmask, nf, 1, NULL, fu,
                       nthreads);
    }
    /* KokkosBlas::scal(x_k, dt, vel) and the vorticity tendency */
    for (int64_t i = 0; i < nv; ++i) {
      for (int c = 0; c < 3; ++c) vk[stage][3 * i + c] = dt * vu[3 * i + c];
      vzk[stage][i] = bve_tendency(vu + 3 * i, dt, Omega);
    }
    for (int64_t i = 0; i < nf; ++i) {
      for (int c = 0; c < 3; ++c) fk[stage][3 * i + c] = dt * fu[3 * i + c];
      fzk[stage][i] = bve_tendency(fu + 3 * i, dt, Omega);
    }
    if (stage < 3) {
      /* KokkosBlas::update(1, x, a, x_k, 0, xwork) */
      const double a = coef[stage];
      for (int64_t i = 0; i < nv; ++i) {
        for (int c = 0; c < 3; ++c)
          vxw[3 * i + c] = vx[3 * i + c] + a * vk[stage][3 * i + c];
        vzw[i] = vzeta[i] + a * vzk[stage][i];
        if (renorm) {
          const double r = sqrt(dot3(vxw + 3 * i, vxw + 3 * i));
          for (int c = 0; c < 3; ++c) vxw[3 * i + c] /= r;
        }
      }
      for (int64_t i = 0; i < nf; ++i) {
        for (int c = 0; c < 3; ++c)
          fxw[3 * i + c] = fx[3 * i + c] + a * fk[stage][3 * i + c];
        fzw[i] = fzeta[i] + a * fzk[stage][i];
        if (renorm) {
          const double r = sqrt(dot3(fxw + 3 * i, fxw + 3 * i));
          for (int c = 0; c < 3; ++c) fxw[3 * i + c] /= r;
        }
      }
    }
  }
  /* BVERK4Update */
  for (int64_t i = 0; i < nv; ++i) {
    for (int c = 0; c < 3; ++c)
      vx[3 * i + c] += sixth * (vk[0][3 * i + c] + vk[3][3 * i + c]) +
                       third * (vk[1][3 * i + c] + vk[2][3 * i + c]);
    vzeta[i] += sixth * (vzk[0][i] + vzk[3][i]) + third * (vzk[1][i] + vzk[2][i]);
    if (renorm) {
      const double r = sqrt(dot3(vx + 3 * i, vx + 3 * i));
      for (int c = 0; c < 3; ++c) vx[3 * i + c] /= r;
    }
  }
  const double* fz3 = compat ? fzk[3] : fzk[2];
  for (int64_t i = 0; i < nf; ++i) {
    for (int c = 0; c < 3; ++c)
      fx[3 * i + c] += sixth * (fk[0][3 * i + c] + fk[3][3 * i + c]) +
                       third * (fk[1][3 * i + c] + fk[2][3 * i + c]);
    fzeta[i] += sixth * (fzk[0][i] + fzk[3][i]) + third * (fzk[1][i] + fz3[i]);
    if (renorm) {
      const double r = sqrt(dot3(fx + 3 * i, fx + 3 * i));
      for (int c = 0; c < 3; ++c) fx[3 * i + c] /= r;
    }
  }
  ora_bve_velocity(vx, nv, fx, fzeta, farea, fmask, nf, 0, NULL, vu, nthreads);
  ora_bve_velocity(fx, nf, fx, fzeta, farea, fmask, nf, 1, NULL, fu, nthreads);

  for (int s = 0; s < 4; ++s) {
    free(vk[s]);
    free(vzk[s]);
    free(fk[s]);
    free(fzk[s]);
  }
  free(vxw);
  free(vzw);
  free(fxw);
  free(fzw);
}

/* ------------------------------------------------------------------ SWE --
 * kzeta_sphere, src/lpm_swe_kernels.hpp:62-75 (accumulates into u).
 * Pinned by: golden vector tests/lpm_swe_kernels_tests.cpp:47-48, tangency. */
void ora_kzeta_sphere(double* u, const double* x, const double* y,
                      double vort, double area, double eps) {
  const double denom = 4 * ORA_PI * (1 - dot3(x, y) + eps * eps);
  const double strength = -vort * area / denom;
  double c[3];
  cross3(c, x, y);
  for (int j = 0; j < 3; ++j) u[j] += c[j] * strength;
}

/* ksigma_sphere, src/lpm_swe_kernels.hpp:86-104: u += P(x) y * (-sigma A) /
 * (4 pi (1 - x.y + eps^2)), P(x) = I - x x^T.
 * Pinned by: golden vector tests/lpm_swe_kernels_tests.cpp:49-50, tangency. */
void ora_ksigma_sphere(double* u, const double* x, const double* y,
                       double div, double area, double eps) {
  const double denom = 4 * ORA_PI * (1 - dot3(x, y) + eps * eps);
  const double strength = -div * area / denom;
  for (int j = 0; j < 3; ++j) {
    double pj = 0;
    for (int k = 0; k < 3; ++k) pj += (((j == k) ? 1.0 : 0.0) - x[j] * x[k]) * y[k];
    u[j] += pj * strength;
  }
}

/* grad_kzeta, src/lpm_swe_kernels.hpp:113-148, restated in closed form:
 *   G = -[ D [y]^T_x + c (y - (1+e) x)^T ] / (4 pi D^2),
 *   c = x cross y, D = 1 - x.y + e, e = eps^2, ([y]^T_x)_ab = eps_abc y_c.
 * G equals minus the Jacobian (w.r.t. x, projected by P(x) on the right) of
 * (x cross y)/(4 pi D).  Row = velocity component, column = derivative.
 * Pinned by: golden vector tests/lpm_swe_kernels_tests.cpp:94-99 (the
 * reference's own test negates it; the function itself matches minus the
 * tabulated values) and by finite differences of ora_kzeta_sphere. */
void ora_grad_kzeta(double* g, const double* x, const double* y, double eps) {
  const double e = eps * eps;
  const double D = 1 - dot3(x, y) + e;
  const double inv = 1.0 / (4 * ORA_PI * D * D);
  double c[3];
  cross3(c, x, y);
  /* eps_abc y_c */
  const double E[3][3] = {{0, y[2], -y[1]}, {-y[2], 0, y[0]}, {y[1], -y[0], 0}};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      g[3 * a + b] = -(D * E[a][b] + c[a] * (y[b] - (1 + e) * x[b])) * inv;
}

/* grad_ksigma, src/lpm_swe_kernels.hpp:178-307, restated in closed form:
 *   G = [ -t D P(x) - D x q^T + q q^T ] / (4 pi D^2),  q = y - t x, t = x.y,
 * the Jacobian (projected by P(x)) of P(x) y / (4 pi D).  Equal to the
 * reference polynomial to rounding on and off the unit sphere.
 * Pinned by: golden vector tests/lpm_swe_kernels_tests.cpp:122-127 and by
 * finite differences of ora_ksigma_sphere. */
void ora_grad_ksigma(double* g, const double* x, const double* y, double eps) {
  const double e = eps * eps;
  const double t = dot3(x, y);
  const double D = 1 - t + e;
  const double inv = 1.0 / (4 * ORA_PI * D * D);
  const double q[3] = {y[0] - t * x[0], y[1] - t * x[1], y[2] - t * x[2]};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      const double P = ((a == b) ? 1.0 : 0.0) - x[a] * x[b];
      g[3 * a + b] = (-t * D * P - D * x[a] * q[b] + q[a] * q[b]) * inv;
    }
}

/* sphere_swe_velocity_sums, src/lpm_swe_kernels.hpp:334-362: 12-tuple of
 * velocity (3) and packed velocity gradient (9) contributed by source y.
 * The reference leaves velz/vels uninitialised before accumulating into them
 * (:341-346); they are zero here. */
void ora_sphere_swe_velocity_sums(double* r, const double* x, const double* y,
                                  double zeta, double sigma, double area,
                                  double eps) {
  double velz[3] = {0, 0, 0}, vels[3] = {0, 0, 0};
  ora_kzeta_sphere(velz, x, y, zeta, area, eps);
  ora_ksigma_sphere(vels, x, y, sigma, area, eps);
  for (int k = 0; k < 3; ++k) r[k] = velz[k] + vels[k];
  double gkz[9], gks[9];
  ora_grad_kzeta(gkz, x, y, eps);
  ora_grad_ksigma(gks, x, y, eps);
  const double rot_str = -zeta * area;
  const double pot_str = -sigma * area;
  for (int k = 0; k < 9; ++k) r[3 + k] = gkz[k] * rot_str + gks[k] * pot_str;
}

/* Gaussian (order-2) PSE Laplacian weight, pse::BivariateOrder2::laplacian
 * (src/lpm_pse.hpp:104-110) as used by LaplacianReducer (:186-202):
 *   (f_j - f_i) A_j * (4/pi) exp(-r^2) / eps^4,  r = |x - y| / eps,
 * restricted to x.y > 0 on the sphere.  The reference's sphere branch feeds
 * PlaneGeometry distances (x, y components only); here r is the 3-D chord.
 * Pinned by: closed-form Laplacians of low-degree polynomials on the sphere
 * (Delta z = -2 z, Delta (3z^2-1) = -6 (3z^2-1)) at O(eps^2). */
double ora_pse_laplacian_pair(const double* x, const double* y, double fy,
                              double fx, double area, double pse_eps) {
  if (dot3(x, y) <= 0) return 0;
  const double d0 = x[0] - y[0], d1 = x[1] - y[1], d2 = x[2] - y[2];
  const double rsq = (d0 * d0 + d1 * d1 + d2 * d2) / (pse_eps * pse_eps);
  const double kern = 4 * exp(-rsq) / ORA_PI;
  const double e2 = pse_eps * pse_eps;
  return (fy - fx) * area * (kern / e2) / e2;
}

/* SphereVertexSums / SphereFaceSums with SphereSweDirectSumReducer,
 * src/lpm_swe_kernels.hpp:578-619, 723-780, 877-930, extended by the PSE
 * surface Laplacian of the sphere SWE right-hand side (sphere_swe_sums_rhs_pse
 * :481-515).  For face targets (collocated != 0) j == i is skipped only when
 * eps == 0, as SphereFaceSums does (:912).  Outputs (nullable): velocity
 * (nt,3), double dot grad(u):grad(u)^T (:773-778), surface Laplacian. */
void ora_swe_sums(const double* tx, const double* ts, int64_t nt,
                  const double* sx, const double* szeta, const double* ssigma,
                  const double* sarea, const uint8_t* smask, const double* ss,
                  int64_t ns, int collocated, double eps, double pse_eps,
                  double* u, double* ddot, double* laps, int nthreads) {
  set_threads(nthreads);
  const int skip_self = collocated && eps == 0;
#pragma omp parallel for schedule(dynamic, 32)
  for (int64_t i = 0; i < nt; ++i) {
    double sums[12] = {0};
    double lap = 0;
    const double* x = tx + 3 * i;
    for (int64_t j = 0; j < ns; ++j) {
      if (smask && smask[j]) continue;
      if (skip_self && i == j) continue;
      double r[12];
      ora_sphere_swe_velocity_sums(r, x, sx + 3 * j, szeta[j], ssigma[j],
                                   sarea[j], eps);
      for (int k = 0; k < 12; ++k) sums[k] += r[k];
      if (laps && pse_eps > 0)
        lap += ora_pse_laplacian_pair(x, sx + 3 * j, ss[j], ts[i], sarea[j],
                                      pse_eps);
    }
    if (u)
      for (int k = 0; k < 3; ++k) u[3 * i + k] = sums[k];
    if (ddot) {
      double dd = 0;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) dd += sums[3 + 3 * a + b] * sums[3 + 3 * b + a];
      ddot[i] = dd;
    }
    if (laps) laps[i] = lap;
  }
}

/* Incompressible2DKernels<SphereGeometry>::kernel_vals summed by
 * Incompressible2DReducer (src/lpm_incompressible2d_kernels.hpp:35-47,
 * 96-136): with arg = 1 - x.y + eps^2 and Gamma = zeta A,
 *   u   += [(x cross y) / (-4 pi arg)] Gamma,
 *   psi += [-log(arg) / (4 pi)] Gamma,
 * masked sources skipped, in source order.  Incompressible2DPassiveSums
 * (collocated = 0, :143-190) never skips j == i; Incompressible2DActiveSums
 * (collocated = 1, :197-247) skips j == i only when eps == 0 (:242).
 * Outputs u (nt, 3) and psi (nt), either may be NULL.
 * Pinned by: at eps = 0 the velocity equals ora_bve_velocity and psi equals
 * ora_bve_stream_fn (restated independently), the RH 5-4 eigenrelation for
 * psi, and a brute-force evaluation of the pair formula with eps > 0. */
void ora_ic2d_sums(const double* tx, int64_t nt, const double* sx,
                   const double* szeta, const double* sarea,
                   const uint8_t* smask, int64_t ns, int collocated,
                   double eps, double* u, double* psi, int nthreads) {
  set_threads(nthreads);
  const int skip_self = collocated && eps == 0.0;
  const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < nt; ++i) {
    double s[4] = {0, 0, 0, 0};
    const double* x = tx + 3 * i;
    for (int64_t j = 0; j < ns; ++j) {
      if (skip_self && i == j) continue;
      if (smask && smask[j]) continue;
      const double* y = sx + 3 * j;
      const double gamma = szeta[j] * sarea[j];
      const double arg = 1 - dot3(x, y) + e2;
      double r[4];
      r[0] = x[1] * y[2] - x[2] * y[1];
      r[1] = x[2] * y[0] - x[0] * y[2];
      r[2] = x[0] * y[1] - x[1] * y[0];
      for (int k = 0; k < 3; ++k) r[k] /= (-4 * ORA_PI * arg);
      r[3] = -log(arg) / (4 * ORA_PI);
      for (int k = 0; k < 4; ++k) s[k] += r[k] * gamma;
    }
    if (u)
      for (int k = 0; k < 3; ++k) u[3 * i + k] = s[k];
    if (psi) psi[i] = s[3];
  }
}

/* One Incompressible2DRK2::advance_timestep_impl on the sphere
 * (src/lpm_incompressible2d_rk2_impl.hpp:75-172, Heun's method) with
 * Incompressible2DTendencies (src/lpm_incompressible2d_kernels.hpp:256-277,
 * dzeta/dt = -2 Omega u_z, CoriolisSphere):
 *   x1 = dt u, z1 = -2 Omega u_z; xw = x + dt u, zw = z + dt z1;
 *   sums at (xw, zw) -> u2; x2 = dt u2, z2 = -2 Omega u2_z;
 *   z += dt/2 (z1 + z2), x += (x1 + x2)/2; sums at the new state.
 * Every stored face is advanced (the reference's ranges are all faces).
 * On entry vu/fu hold the velocity at the current state; on exit the
 * velocity and stream function (vpsi/fpsi) at the new state.
 * Pinned by: exact solid-body rotation to second order, RK2
 * self-convergence order 2. */
void ora_ic2d_rk2_step(double* vx, double* vzeta, double* vu, double* vpsi,
                       int64_t nv, double* fx, double* fzeta, double* fu,
                       double* fpsi, const double* farea, const uint8_t* fmask,
                       int64_t nf, double dt, double Omega, double eps,
                       int nthreads) {
  double* vx1 = calloc(3 * (nv > 0 ? nv : 1), sizeof(double));
  double* vz1 = calloc((nv > 0 ? nv : 1), sizeof(double));
  double* vxw = calloc(3 * (nv > 0 ? nv : 1), sizeof(double));
  double* vzw = calloc((nv > 0 ? nv : 1), sizeof(double));
  double* fx1 = calloc(3 * nf, sizeof(double));
  double* fz1 = calloc(nf, sizeof(double));
  double* fxw = calloc(3 * nf, sizeof(double));
  double* fzw = calloc(nf, sizeof(double));
  for (int64_t i = 0; i < nv; ++i) {
    for (int k = 0; k < 3; ++k) {
      vx1[3 * i + k] = dt * vu[3 * i + k];
      vxw[3 * i + k] = vx[3 * i + k] + dt * vu[3 * i + k];
    }
    vz1[i] = -2.0 * Omega * vu[3 * i + 2];
    vzw[i] = vzeta[i] + dt * vz1[i];
  }
  for (int64_t i = 0; i < nf; ++i) {
    for (int k = 0; k < 3; ++k) {
      fx1[3 * i + k] = dt * fu[3 * i + k];
      fxw[3 * i + k] = fx[3 * i + k] + dt * fu[3 * i + k];
    }
    fz1[i] = -2.0 * Omega * fu[3 * i + 2];
    fzw[i] = fzeta[i] + dt * fz1[i];
  }
  ora_ic2d_sums(vxw, nv, fxw, fzw, farea, fmask, nf, 0, eps, vu, vpsi,
                nthreads);
  ora_ic2d_sums(fxw, nf, fxw, fzw, farea, fmask, nf, 1, eps, fu, fpsi,
                nthreads);
  for (int64_t i = 0; i < nv; ++i) {
    const double z2 = -2.0 * Omega * vu[3 * i + 2];
    vzeta[i] += 0.5 * dt * vz1[i] + 0.5 * dt * z2;
    for (int k = 0; k < 3; ++k)
      vx[3 * i + k] += 0.5 * vx1[3 * i + k] + 0.5 * (dt * vu[3 * i + k]);
  }
  for (int64_t i = 0; i < nf; ++i) {
    const double z2 = -2.0 * Omega * fu[3 * i + 2];
    fzeta[i] += 0.5 * dt * fz1[i] + 0.5 * dt * z2;
    for (int k = 0; k < 3; ++k)
      fx[3 * i + k] += 0.5 * fx1[3 * i + k] + 0.5 * (dt * fu[3 * i + k]);
  }
  ora_ic2d_sums(vx, nv, fx, fzeta, farea, fmask, nf, 0, eps, vu, vpsi,
                nthreads);
  ora_ic2d_sums(fx, nf, fx, fzeta, farea, fmask, nf, 1, eps, fu, fpsi,
                nthreads);
  free(vx1);
  free(vz1);
  free(vxw);
  free(vzw);
  free(fx1);
  free(fz1);
  free(fxw);
  free(fzw);
}

/* CoriolisSphere, src/lpm_coriolis.hpp:154-195. */
static inline double cor_f(const double* x, double Om) { return 2 * Om * x[2]; }
static inline double cor_dfdt(const double* u, double Om) { return 2 * Om * u[2]; }
static inline double cor_grad_f_cross_u(const double* x, const double* u,
                                        double Om) {
  return -2 * Om * (-u[0] * x[1] + u[1] * x[0]);
}

/* SWEVorticityDivergenceHeightTendencies / ...AreaTendencies (sphere
 * branch), src/lpm_swe_kernels.hpp:941-999, 1010-1069.  kind 0: third output
 * is the depth tendency -sigma h dt; kind 1: the area tendency sigma A dt. */
void ora_swe_tendencies(const double* x, const double* u, const double* zeta,
                        const double* sigma, const double* h_or_a,
                        const double* ddot, const double* laps, int64_t n,
                        double Omega, double g, double dt, int kind,
                        double* dzeta, double* dsigma, double* dh_or_da) {
  for (int64_t i = 0; i < n; ++i) {
    const double* xi = x + 3 * i;
    const double* ui = u + 3 * i;
    const double f = cor_f(xi, Omega);
    dzeta[i] = (-cor_dfdt(ui, Omega) - (zeta[i] + f) * sigma[i]) * dt;
    dsigma[i] = (f * zeta[i] + cor_grad_f_cross_u(xi, ui, Omega) - ddot[i] -
                 g * laps[i] - dot3(ui, ui)) * dt;
    dh_or_da[i] = kind == 0 ? (-sigma[i] * h_or_a[i]) * dt
                            : (sigma[i] * h_or_a[i]) * dt;
  }
}

/* One sphere SWE RK4 step in the stage layout of SWERK4::advance_timestep
 * (src/lpm_swe_rk4_impl.hpp:204-445) with the sphere sums, sphere
 * tendencies and zero bottom topography.  Passive state: vx, vzeta, vsigma,
 * vh (+ velocity vu, double dot vdd, surface Laplacian vlap, surface vs).
 * Active state: fx, fzeta, fsigma, farea (+ fmass fixed, fu, fdd, flap, fs,
 * fh derived).  On entry u, ddot, laps hold the sums at the current state.
 * compat != 0 keeps the reference's stage-4 position increment at zero (the
 * reference never assigns passive_x4/active_x4, :348-400); compat == 0 uses
 * dt*u after the stage-4 sums.  renorm != 0 projects every work position
 * and the updated positions onto the unit sphere (not done by the
 * reference).  Pinned by: steady TC2 state (closed forms), fourth-order
 * self-convergence on a perturbed TC2 state. */
static void project_rows(double* x, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    const double r = sqrt(dot3(x + 3 * i, x + 3 * i));
    for (int c = 0; c < 3; ++c) x[3 * i + c] /= r;
  }
}

void ora_swe_rk4_step(double* vx, double* vzeta, double* vsigma, double* vh,
                      double* vu, double* vdd, double* vlap, double* vs,
                      int64_t nv, double* fx, double* fzeta, double* fsigma,
                      double* farea, const double* fmass, double* fh,
                      double* fu, double* fdd, double* flap, double* fs,
                      const uint8_t* fmask, int64_t nf, double dt,
                      double Omega, double g, double eps, double pse_eps,
                      int compat, int renorm, int nthreads) {
  const double sixth = 1.0 / 6.0, third = 1.0 / 3.0;
  const int64_t nvv = nv > 0 ? nv : 1, nff = nf > 0 ? nf : 1;
  double *vk[4], *vzk[4], *vsk[4], *vhk[4], *fk[4], *fzk[4], *fsk[4], *fak[4];
  for (int s = 0; s < 4; ++s) {
    vk[s] = (double*)calloc(3 * nvv, sizeof(double));
    vzk[s] = (double*)calloc(nvv, sizeof(double));
    vsk[s] = (double*)calloc(nvv, sizeof(double));
    vhk[s] = (double*)calloc(nvv, sizeof(double));
    fk[s] = (double*)calloc(3 * nff, sizeof(double));
    fzk[s] = (double*)calloc(nff, sizeof(double));
    fsk[s] = (double*)calloc(nff, sizeof(double));
    fak[s] = (double*)calloc(nff, sizeof(double));
  }
  double* vxw = (double*)calloc(3 * nvv, sizeof(double));
  double* vzw = (double*)calloc(nvv, sizeof(double));
  double* vsw = (double*)calloc(nvv, sizeof(double));
  double* vhw = (double*)calloc(nvv, sizeof(double));
  double* fxw = (double*)calloc(3 * nff, sizeof(double));
  double* fzw = (double*)calloc(nff, sizeof(double));
  double* fsw = (double*)calloc(nff, sizeof(double));
  double* faw = (double*)calloc(nff, sizeof(double));
  const double coef[3] = {0.5, 0.5, 1.0};

  /* stage 1 tendencies at the current state */
  for (int64_t i = 0; i < nv; ++i)
    for (int c = 0; c < 3; ++c) vk[0][3 * i + c] = dt * vu[3 * i + c];
  for (int64_t i = 0; i < nf; ++i)
    for (int c = 0; c < 3; ++c) fk[0][3 * i + c] = dt * fu[3 * i + c];
  ora_swe_tendencies(vx, vu, vzeta, vsigma, vh, vdd, vlap, nv, Omega, g, dt, 0,
                     vzk[0], vsk[0], vhk[0]);
  ora_swe_tendencies(fx, fu, fzeta, fsigma, farea, fdd, flap, nf, Omega, g, dt,
                     1, fzk[0], fsk[0], fak[0]);

  for (int stage = 1; stage < 4; ++stage) {
    const double a = coef[stage - 1];
    if (stage > 1) {
      for (int64_t i = 0; i < nv; ++i)
        for (int c = 0; c < 3; ++c) vk[stage - 1][3 * i + c] = dt * vu[3 * i + c];
      for (int64_t i = 0; i < nf; ++i)
        for (int c = 0; c < 3; ++c) fk[stage - 1][3 * i + c] = dt * fu[3 * i + c];
    }
    for (int64_t i = 0; i < nv; ++i) {
      for (int c = 0; c < 3; ++c)
        vxw[3 * i + c] = vx[3 * i + c] + a * vk[stage - 1][3 * i + c];
      vzw[i] = vzeta[i] + a * vzk[stage - 1][i];
      vsw[i] = vsigma[i] + a * vsk[stage - 1][i];
      vhw[i] = vh[i] + a * vhk[stage - 1][i];
      vs[i] = vhw[i]; /* SetSurfaceFromDepth, zero bottom */
    }
    for (int64_t i = 0; i < nf; ++i) {
      for (int c = 0; c < 3; ++c)
        fxw[3 * i + c] = fx[3 * i + c] + a * fk[stage - 1][3 * i + c];
      fzw[i] = fzeta[i] + a * fzk[stage - 1][i];
      fsw[i] = fsigma[i] + a * fsk[stage - 1][i];
      faw[i] = farea[i] + a * fak[stage - 1][i];
      if (!(fmask && fmask[i])) { /* SetDepthAndSurfaceFromMassAndArea */
        fh[i] = fmass[i] / faw[i];
        fs[i] = fh[i];
      }
    }
    if (renorm) {
      project_rows(vxw, nv);
      project_rows(fxw, nf);
    }
    ora_swe_sums(vxw, vs, nv, fxw, fzw, fsw, faw, fmask, fs, nf, 0, eps,
                 pse_eps, vu, vdd, vlap, nthreads);
    ora_swe_sums(fxw, fs, nf, fxw, fzw, fsw, faw, fmask, fs, nf, 1, eps,
                 pse_eps, fu, fdd, flap, nthreads);
    ora_swe_tendencies(vxw, vu, vzw, vsw, vhw, vdd, vlap, nv, Omega, g, dt, 0,
                       vzk[stage], vsk[stage], vhk[stage]);
    ora_swe_tendencies(fxw, fu, fzw, fsw, faw, fdd, flap, nf, Omega, g, dt, 1,
                       fzk[stage], fsk[stage], fak[stage]);
  }
  if (!compat) {
    for (int64_t i = 0; i < nv; ++i)
      for (int c = 0; c < 3; ++c) vk[3][3 * i + c] = dt * vu[3 * i + c];
    for (int64_t i = 0; i < nf; ++i)
      for (int c = 0; c < 3; ++c) fk[3][3 * i + c] = dt * fu[3 * i + c];
  }
  /* SWERK4Update */
  for (int64_t i = 0; i < nv; ++i) {
    for (int c = 0; c < 3; ++c)
      vx[3 * i + c] += sixth * (vk[0][3 * i + c] + vk[3][3 * i + c]) +
                       third * (vk[1][3 * i + c] + vk[2][3 * i + c]);
    vzeta[i] += sixth * (vzk[0][i] + vzk[3][i]) + third * (vzk[1][i] + vzk[2][i]);
    vsigma[i] += sixth * (vsk[0][i] + vsk[3][i]) + third * (vsk[1][i] + vsk[2][i]);
    vh[i] += sixth * (vhk[0][i] + vhk[3][i]) + third * (vhk[1][i] + vhk[2][i]);
    vs[i] = vh[i];
  }
  for (int64_t i = 0; i < nf; ++i) {
    for (int c = 0; c < 3; ++c)
      fx[3 * i + c] += sixth * (fk[0][3 * i + c] + fk[3][3 * i + c]) +
                       third * (fk[1][3 * i + c] + fk[2][3 * i + c]);
    fzeta[i] += sixth * (fzk[0][i] + fzk[3][i]) + third * (fzk[1][i] + fzk[2][i]);
    fsigma[i] += sixth * (fsk[0][i] + fsk[3][i]) + third * (fsk[1][i] + fsk[2][i]);
    farea[i] += sixth * (fak[0][i] + fak[3][i]) + third * (fak[1][i] + fak[2][i]);
    if (!(fmask && fmask[i])) {
      fh[i] = fmass[i] / farea[i];
      fs[i] = fh[i];
    }
  }
  if (renorm) {
    project_rows(vx, nv);
    project_rows(fx, nf);
  }
  ora_swe_sums(vx, vs, nv, fx, fzeta, fsigma, farea, fmask, fs, nf, 0, eps,
               pse_eps, vu, vdd, vlap, nthreads);
  ora_swe_sums(fx, fs, nf, fx, fzeta, fsigma, farea, fmask, fs, nf, 1, eps,
               pse_eps, fu, fdd, flap, nthreads);

  for (int s = 0; s < 4; ++s) {
    free(vk[s]); free(vzk[s]); free(vsk[s]); free(vhk[s]);
    free(fk[s]); free(fzk[s]); free(fsk[s]); free(fak[s]);
  }
  free(vxw); free(vzw); free(vsw); free(vhw);
  free(fxw); free(fzw); free(fsw); free(faw);
}

/* One sphere SWE RK2 (Heun) step in the stage layout of
 * SWERK2::advance_timestep_impl (src/lpm_swe_rk2_impl.hpp:80-257):
 * increments at the state (positions dt u, tendencies), the work state
 * state + increments (:104-121) with its surface (:122-131), the sums at the
 * work state (:155-166) and increments there (:168-185), then
 * state += increments1 / 2 + increments2 / 2 (:190-205), the surface and the
 * sums at the new state (:207-232).  The surface Laplacian is the PSE sum
 * (the reference evaluates a GMLS operator there).  Arguments as
 * ora_swe_rk4_step (renorm: work and updated positions projected onto the
 * unit sphere).  Pinned by: steady TC2 state, second-order agreement with
 * the RK4 step (error ratio under dt halving), and, for dt -> 0, the stage
 * layout's first-order term. */
void ora_swe_rk2_step(double* vx, double* vzeta, double* vsigma, double* vh,
                      double* vu, double* vdd, double* vlap, double* vs,
                      int64_t nv, double* fx, double* fzeta, double* fsigma,
                      double* farea, const double* fmass, double* fh,
                      double* fu, double* fdd, double* flap, double* fs,
                      const uint8_t* fmask, int64_t nf, double dt,
                      double Omega, double g, double eps, double pse_eps,
                      int renorm, int nthreads) {
  const int64_t nvv = nv > 0 ? nv : 1, nff = nf > 0 ? nf : 1;
  double* vx1 = (double*)calloc(3 * nvv, sizeof(double));
  double* vx2 = (double*)calloc(3 * nvv, sizeof(double));
  double* vxw = (double*)calloc(3 * nvv, sizeof(double));
  double* fx1 = (double*)calloc(3 * nff, sizeof(double));
  double* fx2 = (double*)calloc(3 * nff, sizeof(double));
  double* fxw = (double*)calloc(3 * nff, sizeof(double));
  /* per kind: zeta, sigma, h|A increments of stages 1, 2 and work values */
  double* v1 = (double*)calloc(3 * nvv, sizeof(double));
  double* v2 = (double*)calloc(3 * nvv, sizeof(double));
  double* vw = (double*)calloc(3 * nvv, sizeof(double));
  double* f1 = (double*)calloc(3 * nff, sizeof(double));
  double* f2 = (double*)calloc(3 * nff, sizeof(double));
  double* fw = (double*)calloc(3 * nff, sizeof(double));

  /* stage 1 */
  for (int64_t i = 0; i < 3 * nv; ++i) vx1[i] = dt * vu[i];
  for (int64_t i = 0; i < 3 * nf; ++i) fx1[i] = dt * fu[i];
  ora_swe_tendencies(vx, vu, vzeta, vsigma, vh, vdd, vlap, nv, Omega, g, dt, 0,
                     v1, v1 + nvv, v1 + 2 * nvv);
  ora_swe_tendencies(fx, fu, fzeta, fsigma, farea, fdd, flap, nf, Omega, g, dt,
                     1, f1, f1 + nff, f1 + 2 * nff);
  /* work state at t + dt */
  for (int64_t i = 0; i < nv; ++i) {
    for (int c = 0; c < 3; ++c) vxw[3 * i + c] = vx[3 * i + c] + vx1[3 * i + c];
    vw[i] = vzeta[i] + v1[i];
    vw[nvv + i] = vsigma[i] + v1[nvv + i];
    vw[2 * nvv + i] = vh[i] + v1[2 * nvv + i];
    vs[i] = vw[2 * nvv + i];
  }
  for (int64_t i = 0; i < nf; ++i) {
    for (int c = 0; c < 3; ++c) fxw[3 * i + c] = fx[3 * i + c] + fx1[3 * i + c];
    fw[i] = fzeta[i] + f1[i];
    fw[nff + i] = fsigma[i] + f1[nff + i];
    fw[2 * nff + i] = farea[i] + f1[2 * nff + i];
    if (!(fmask && fmask[i])) {
      fh[i] = fmass[i] / fw[2 * nff + i];
      fs[i] = fh[i];
    }
  }
  if (renorm) {
    project_rows(vxw, nv);
    project_rows(fxw, nf);
  }
  /* stage 2 */
  ora_swe_sums(vxw, vs, nv, fxw, fw, fw + nff, fw + 2 * nff, fmask, fs, nf, 0,
               eps, pse_eps, vu, vdd, vlap, nthreads);
  ora_swe_sums(fxw, fs, nf, fxw, fw, fw + nff, fw + 2 * nff, fmask, fs, nf, 1,
               eps, pse_eps, fu, fdd, flap, nthreads);
  ora_swe_tendencies(vxw, vu, vw, vw + nvv, vw + 2 * nvv, vdd, vlap, nv, Omega,
                     g, dt, 0, v2, v2 + nvv, v2 + 2 * nvv);
  ora_swe_tendencies(fxw, fu, fw, fw + nff, fw + 2 * nff, fdd, flap, nf, Omega,
                     g, dt, 1, f2, f2 + nff, f2 + 2 * nff);
  for (int64_t i = 0; i < 3 * nv; ++i) vx2[i] = dt * vu[i];
  for (int64_t i = 0; i < 3 * nf; ++i) fx2[i] = dt * fu[i];
  /* update */
  for (int64_t i = 0; i < nv; ++i) {
    for (int c = 0; c < 3; ++c)
      vx[3 * i + c] += 0.5 * vx1[3 * i + c] + 0.5 * vx2[3 * i + c];
    vzeta[i] += 0.5 * v1[i] + 0.5 * v2[i];
    vsigma[i] += 0.5 * v1[nvv + i] + 0.5 * v2[nvv + i];
    vh[i] += 0.5 * v1[2 * nvv + i] + 0.5 * v2[2 * nvv + i];
    vs[i] = vh[i];
  }
  for (int64_t i = 0; i < nf; ++i) {
    for (int c = 0; c < 3; ++c)
      fx[3 * i + c] += 0.5 * fx1[3 * i + c] + 0.5 * fx2[3 * i + c];
    fzeta[i] += 0.5 * f1[i] + 0.5 * f2[i];
    fsigma[i] += 0.5 * f1[nff + i] + 0.5 * f2[nff + i];
    farea[i] += 0.5 * f1[2 * nff + i] + 0.5 * f2[2 * nff + i];
    if (!(fmask && fmask[i])) {
      fh[i] = fmass[i] / farea[i];
      fs[i] = fh[i];
    }
  }
  if (renorm) {
    project_rows(vx, nv);
    project_rows(fx, nf);
  }
  ora_swe_sums(vx, vs, nv, fx, fzeta, fsigma, farea, fmask, fs, nf, 0, eps,
               pse_eps, vu, vdd, vlap, nthreads);
  ora_swe_sums(fx, fs, nf, fx, fzeta, fsigma, farea, fmask, fs, nf, 1, eps,
               pse_eps, fu, fdd, flap, nthreads);
  free(vx1); free(vx2); free(vxw); free(fx1); free(fx2); free(fxw);
  free(v1); free(v2); free(vw); free(f1); free(f2); free(fw);
}

/* ------------------------------------------------------------ galleries --
 * atan4: longitude in [0, 2 pi) (src/util/lpm_math.hpp:70-104). */
static double atan4(double y, double x) {
  const double tol = 2.220446049250313e-16;
  if (fabs(x) < tol) {
    if (y > 0.0) return 0.5 * ORA_PI;
    if (y < 0.0) return 1.5 * ORA_PI;
    return 0.0;
  }
  if (fabs(y) < tol) return x > 0.0 ? 0.0 : ORA_PI;
  const double th = atan2(fabs(y), fabs(x));
  if (x > 0.0 && y > 0.0) return th;
  if (x < 0.0 && y > 0.0) return ORA_PI - th;
  if (x < 0.0 && y < 0.0) return ORA_PI + th;
  return 2 * ORA_PI - th;
}

/* RossbyHaurwitz54, src/lpm_vorticity_gallery.hpp:104-147:
 *   zeta = 2 u0 z + 30 amp cos(4 lon) z (z^2 - 1)^2. */
void ora_rh54_vorticity(const double* x, int64_t n, double u0, double amp,
                        double* zeta) {
  for (int64_t i = 0; i < n; ++i) {
    const double* p = x + 3 * i;
    const double lon = atan4(p[1], p[0]);
    const double z = p[2];
    const double P54 = z * (z * z - 1) * (z * z - 1);
    zeta[i] = 2 * u0 * z + 30 * amp * cos(4 * lon) * P54;
  }
}

/* RossbyWave54Velocity, src/lpm_velocity_gallery.hpp:250-288 (t = 0). */
void ora_rh54_velocity(const double* x, int64_t n, double u0, double amp,
                       double* u) {
  for (int64_t i = 0; i < n; ++i) {
    const double* p = x + 3 * i;
    const double lat = atan2(p[2], sqrt(p[0] * p[0] + p[1] * p[1]));
    const double lon = atan4(p[1], p[0]);
    const double cl = cos(lat);
    const double uu = amp * 0.5 * cos(4 * lon) * cl * cl * cl * (5 * cos(2 * lat) - 3);
    const double vv = amp * 4 * cl * cl * cl * sin(lat) * sin(4 * lon);
    u[3 * i + 0] = -u0 * p[1] - uu * sin(lon) - vv * sin(lat) * cos(lon);
    u[3 * i + 1] = u0 * p[0] + uu * cos(lon) - vv * sin(lat) * sin(lon);
    u[3 * i + 2] = vv * cos(lat);
  }
}

/* GaussianVortexSphere, src/lpm_vorticity_gallery.hpp:59-102:
 *   zeta = str exp(-b^2 (1 - x.x0)) - gauss_const. */
void ora_gaussian_vorticity(const double* x, int64_t n, const double* ctr,
                            double strength, double b, double gauss_const,
                            double* zeta) {
  for (int64_t i = 0; i < n; ++i) {
    const double* p = x + 3 * i;
    const double distsq = 1.0 - p[0] * ctr[0] - p[1] * ctr[1] - p[2] * ctr[2];
    zeta[i] = strength * exp(-b * b * distsq) - gauss_const;
  }
}

/* ErrNorms, src/lpm_error.hpp:81-131, src/lpm_error_impl.hpp:61-111,155-182:
 * area-weighted relative l1, l2, linf of |appx - exact| (vector magnitude
 * when dim == 3), with FloatingPoint::safe_denominator. */
void ora_err_norms(const double* appx, const double* exact, const double* wt,
                   int64_t n, int dim, double* out) {
  double l1n = 0, l1d = 0, l2n = 0, l2d = 0, lin = 0, lid = 0;
  for (int64_t i = 0; i < n; ++i) {
    double em = 0, xm = 0;
    for (int c = 0; c < dim; ++c) {
      const double e = appx[dim * i + c] - exact[dim * i + c];
      em += e * e;
      xm += exact[dim * i + c] * exact[dim * i + c];
    }
    em = sqrt(em);
    xm = sqrt(xm);
    const double w = wt ? wt[i] : 1.0;
    l1n += em * w;
    l1d += xm * w;
    l2n += em * em * w;
    l2d += xm * xm * w;
    if (em > lin) lin = em;
    if (xm > lid) lid = xm;
  }
  const double tol = 2.220446049250313e-16;
  out[0] = l1n * (l1d / (l1d * l1d + tol * tol));
  out[1] = sqrt(l2n * (l2d / (l2d * l2d + tol * tol)));
  out[2] = lin * (lid / (lid * lid + tol * tol));
}
