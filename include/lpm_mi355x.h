/*
 * lpm_mi355x.h — C ABI of the MI355X (gfx950) direct-sum library.
 *
 * This is the drop-in boundary for LPM's all-pairs spherical direct sums and
 * the RK4 steppers that drive them.  Every entry point takes plain pointers
 * and sizes; no C++, Kokkos, HIP or torch types cross it.  Each declaration
 * names the reference interface it replaces (pbosler/lpm file:line).
 *
 * Conventions
 *  - Device pointers: every array argument is device memory unless the
 *    parameter name ends in _host.
 *  - Coordinates / vectors: element (i, k), k in {0,1,2}, lives at
 *    base[i * es + k * cs].  A Kokkos View<double*[3]> maps without copies:
 *    LayoutLeft (GPU backends) es = 1, cs = view.stride(1);
 *    LayoutRight (OpenMP) es = 3, cs = 1.
 *  - Masks are one byte per face (Kokkos View<bool*>): nonzero = divided face,
 *    excluded as a source.  A NULL mask means no face is divided.
 *  - Streams are hipStream_t passed as void* (NULL = the default stream).
 *    Calls are stream-ordered and asynchronous unless stated otherwise.
 *  - Return value: LPM_OK (0) or a negative lpm_status; the message is
 *    available from lpm_last_error(ctx).  Nothing throws across the ABI.
 *  - A context is single-threaded (not reentrant); distinct contexts may be
 *    used from distinct host threads on distinct streams.
 *  - A context's workspaces (packed sources, mask scan, RK stage buffers,
 *    work queues) are shared by all its calls.  When a call arrives on a
 *    different stream than the context's previous call, the library orders
 *    it after the previous stream's work (an event wait, no host sync).  A
 *    call on a stream that is being captured by the caller is not ordered
 *    this way: order such streams yourself.
 *  - Sums whose target blocks would leave the device underfilled split each
 *    block's sources into chunks and add the chunk partials in a second
 *    kernel (the stream function and Incompressible2D sums always when
 *    they have fewer target blocks than eight rounds of the blocks the
 *    device holds at once; BVE velocities through the *_splitK variants).
 *    The partials live in a context workspace grown on demand; a call that
 *    cannot grow it (on a stream the caller is capturing, or out of device
 *    memory) runs unsplit.  The chunk count follows from the problem size
 *    and the device, so results are reproducible on one device.
 */
#ifndef LPM_MI355X_H
#define LPM_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LPM_ABI_VERSION 1

typedef enum lpm_status {
  LPM_OK = 0,
  LPM_ERR_ARG = -1,       /* invalid argument (null pointer, size, stride) */
  LPM_ERR_HIP = -2,       /* HIP runtime error */
  LPM_ERR_NONFINITE = -3, /* LPM_FLAG_CHECK_FINITE found a NaN/Inf output */
  LPM_ERR_STATE = -4,     /* call out of order / incompatible context state */
  LPM_ERR_NOMEM = -5,     /* device allocation failed */
  LPM_ERR_INTERNAL = -6
} lpm_status;

/* LPM_MIXED evaluates 1 - x.y as |x - y|^2 / 2, which holds for positions
 * on the unit sphere: off it by delta = |x|^2 - 1 the kernel is off by
 * about delta / 2.  The time steppers keep positions on the sphere only
 * with LPM_RK4_RENORMALIZE; without it, stage positions leave the sphere by
 * O((dt |u|)^2) and the mixed sums then differ from FP64 by that much
 * (1e-5 relative at dt |u| = 0.007), not by the ~1e-7 of the pair kernel.
 * The C++ steppers (lpm/*.hpp) print a warning to stderr when they are
 * constructed with LPM_MIXED and without LPM_RK4_RENORMALIZE. */
typedef enum lpm_precision {
  LPM_FP64 = 0,  /* FP64 pair arithmetic, the reference's formula order */
  LPM_MIXED = 1  /* FP32 pair kernel (|x-y|^2/2 form), FP64 carries/state */
} lpm_precision;

/* Flags (bitwise or). */
#define LPM_FLAG_CHECK_FINITE 0x1u   /* scan outputs, return LPM_ERR_NONFINITE */
#define LPM_FLAG_SKIP_MASKED_TARGETS 0x2u /* collocated: leave divided-face
                                            targets untouched (not computed) */
#define LPM_RK4_REFERENCE_COMPAT 0x10u /* reproduce the reference's stage
                                          arithmetic (see lpm_bve_rk4_step) */
#define LPM_RK4_RENORMALIZE 0x20u /* project stage and final positions onto
                                     the unit sphere (FP64) */
#define LPM_RK4_USE_GRAPH 0x40u   /* capture the step in a hipGraph once and
                                     replay it on later identical calls */
/* Captured step graphs a context keeps (keyed by fields, dt, Omega, flags,
 * stream and the kernel-variant choice: lpm_set_variant, lpm_tune_load and
 * an autotune make the next graph step capture again); beyond this the
 * least recently replayed one is destroyed. */
#define LPM_GRAPH_CACHE_MAX 16

typedef struct lpm_ctx lpm_ctx;

/* ------------------------------------------------------------ context ---- */
int lpm_abi_version(void);

/* Creates a context bound to HIP device `device` (the device is made current
 * on the calling thread).  No reference counterpart: replaces the implicit
 * Kokkos::DefaultExecutionSpace (src/lpm_kokkos_defs.hpp:18). */
int lpm_ctx_create(int device, lpm_ctx** out);
int lpm_ctx_destroy(lpm_ctx* ctx);

/* Last error message of `ctx` (or of the calling thread when ctx is NULL,
 * e.g. after a failed lpm_ctx_create).  Never NULL. */
const char* lpm_last_error(const lpm_ctx* ctx);

/* Blocks until all work the context issued on `stream` has finished. */
int lpm_stream_synchronize(lpm_ctx* ctx, void* stream);

/* Number of captured step graphs the context holds (LPM_RK4_USE_GRAPH),
 * at most LPM_GRAPH_CACHE_MAX.  No reference counterpart. */
int lpm_graph_cache_size(const lpm_ctx* ctx, int* n_graphs);

/* -------------------------------------------------------------- mesh ---- */
/* Uniformly refined spherical quadtree mesh (host memory).  Replaces
 * MeshSeed<Seed> + PolyMesh2d<Seed>::tree_init
 * (src/mesh/lpm_mesh_seed.cpp:20-304, src/mesh/lpm_polymesh2d_impl.hpp:25-42)
 * as the producer of the direct sum's inputs; particles are created in the
 * reference's order. */
typedef struct lpm_mesh lpm_mesh;
#define LPM_SEED_ICOS_TRI 0   /* IcosTriSphereSeed */
#define LPM_SEED_CUBED_SPHERE 1 /* CubedSphereSeed */

int lpm_mesh_create(int seed, int depth, lpm_mesh** out);
int lpm_mesh_destroy(lpm_mesh* mesh);
/* Counts: vertices, stored faces (leaves + divided), leaf faces, edges. */
int lpm_mesh_sizes(const lpm_mesh* mesh, int64_t* n_verts, int64_t* n_faces,
                   int64_t* n_leaves, int64_t* n_edges);
/* MeshSeed::set_max_allocations (src/mesh/lpm_mesh_seed.cpp:266-279). */
int lpm_mesh_max_allocations(int seed, int depth, int64_t* n_verts,
                             int64_t* n_edges, int64_t* n_faces);
/* Copies into host arrays; xyz arrays are SoA [3][n] (x row, y row, z row).
 * Any output may be NULL. */
int lpm_mesh_copy_vertices(const lpm_mesh* mesh, double* xyz_host);
int lpm_mesh_copy_faces(const lpm_mesh* mesh, double* xyz_host,
                        double* area_host, uint8_t* mask_host,
                        int32_t* parent_host);
/* Face corner vertex indices, [n_faces][verts_per_face] row-major. */
int lpm_mesh_copy_face_vertices(const lpm_mesh* mesh, int32_t* verts_host,
                                int* verts_per_face);
/* Face tree levels, 0 for seed faces (Faces::level,
 * src/mesh/lpm_faces.hpp:147). */
int lpm_mesh_copy_face_levels(const lpm_mesh* mesh, int32_t* level_host);
/* Adaptive refinement, PolyMesh2d::divide_flagged_faces
 * (src/mesh/lpm_polymesh2d_impl.hpp:129-173): in index order, every face
 * i < n_faces with flags_host[i] != 0 whose level is <= depth + amr_limit is
 * divided (a flagged face that is already divided stays as it is,
 * PolyMesh2d::divide_face :46-53), new vertices and faces are appended and
 * the leaf count is updated.  n_flags must equal n_faces.  max_faces > 0
 * plays the reference's allocation limit: if the flag count exceeds
 * (max_faces - n_faces) / 4 nothing is divided and *n_divided = -1.
 * Otherwise *n_divided is the reference's refine count (flagged faces within
 * the level limit) and *limit_reached != 0 if a flagged face was beyond it.
 * Copy the arrays again afterwards (sizes grow). */
int lpm_mesh_divide_flagged_faces(lpm_mesh* mesh, const uint8_t* flags_host,
                                  int64_t n_flags, int amr_limit,
                                  int64_t max_faces, int64_t* n_divided,
                                  int* limit_reached);

/* --------------------------------------------------------- BVE sums ---- */
/* Velocity of the barotropic vorticity equation on the unit sphere by the
 * all-pairs Biot-Savart sum over the undivided faces:
 *   u(x_i) = sum_j (x_i cross y_j) (-zeta_j A_j) / (4 pi (1 - x_i.y_j + eps^2))
 * Replaces BVEVertexVelocity (collocated = 0) and BVEFaceVelocity
 * (collocated = 1) with VelocityReduceDistinct / VelocityReduceCollocated and
 * biot_savart (src/lpm_bve_sphere_kernels.hpp:53-82, 179-211, 249-274,
 * 365-394; src/lpm_sphere_functions.hpp:44-57).
 *   tgt_*: target positions; n_tgt targets.  With collocated = 1 the targets
 *          are the sources (pass the same array, n_tgt = n_src) and j == i
 *          is skipped; with LPM_FLAG_SKIP_MASKED_TARGETS, divided-face
 *          targets are not computed and their output is left untouched.
 *   src_*: face centres, relative vorticity, area, mask; n_src stored faces
 *          (the reference's n_faces_host(), divided faces included).
 *   eps:   kernel smoothing (the reference uses 0).
 *   out_u: velocity at the targets (may alias nothing else). */
int lpm_bve_velocity(lpm_ctx* ctx, const double* tgt_x, int64_t tgt_es,
                     int64_t tgt_cs, int64_t n_tgt, const double* src_x,
                     int64_t src_es, int64_t src_cs, const double* src_zeta,
                     const double* src_area, const uint8_t* src_mask,
                     int64_t n_src, int collocated, double eps, int precision,
                     unsigned flags, double* out_u, int64_t out_es,
                     int64_t out_cs, void* stream);

/* Allocation-free form of lpm_bve_velocity: same arguments and the same
 * kernel variant, so the output is bit-equal to lpm_bve_velocity's.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP stream capture, tight serving loops).
 * Every workspace (mask scan, packed sources, finite flag) is carved from
 * one static device array of at most 56 MiB per device, held until the
 * process ends; the library looks up each device's copy when a context is
 * first created on it.  The call makes no hipMalloc or hipFree.
 * n_src may not exceed lpm_bve_velocity_noalloc_max_sources() (524288
 * stored faces): larger calls return LPM_ERR_ARG.  All calls on one device
 * share the array and it keeps the last call's contents (each call rewrites
 * everything it reads): calls that overlap on one GPU race, so issue them
 * on one stream or order them.  LPM_FLAG_CHECK_FINITE synchronises the
 * stream (not usable under capture). */
int lpm_bve_velocity_noalloc(lpm_ctx* ctx, const double* tgt_x,
                             int64_t tgt_es, int64_t tgt_cs, int64_t n_tgt,
                             const double* src_x, int64_t src_es,
                             int64_t src_cs, const double* src_zeta,
                             const double* src_area, const uint8_t* src_mask,
                             int64_t n_src, int collocated, double eps,
                             int precision, unsigned flags, double* out_u,
                             int64_t out_es, int64_t out_cs, void* stream);
/* Largest n_src lpm_bve_velocity_noalloc accepts. */
int64_t lpm_bve_velocity_noalloc_max_sources(void);

/* Stream function psi(x_i) = sum_j log(1 - x_i.y_j) (-zeta_j A_j) / (4 pi).
 * Replaces BVEVertexStreamFn / BVEFaceStreamFn with StreamReduceDistinct /
 * StreamReduceCollocated and greens_fn (src/lpm_bve_sphere_kernels.hpp:20-48,
 * 141-170, 218-242, 329-356; src/lpm_sphere_functions.hpp:21-29).
 * Arguments as lpm_bve_velocity; eps must be 0. */
int lpm_bve_stream_fn(lpm_ctx* ctx, const double* tgt_x, int64_t tgt_es,
                      int64_t tgt_cs, int64_t n_tgt, const double* src_x,
                      int64_t src_es, int64_t src_cs, const double* src_zeta,
                      const double* src_area, const uint8_t* src_mask,
                      int64_t n_src, int collocated, int precision,
                      unsigned flags, double* out_psi, void* stream);

/* Stream function and velocity in one pass (the pair formulas of
 * lpm_bve_stream_fn and lpm_bve_velocity; both outputs required unless
 * NULL).  Replaces BVEVertexSolve (collocated = 0) and BVEFaceSolve
 * (collocated = 1, j == i skipped), which run StreamReduceDistinct /
 * StreamReduceCollocated and VelocityReduceDistinct /
 * VelocityReduceCollocated as two reductions per target
 * (src/lpm_bve_sphere_kernels.hpp:84-132, 276-320); here they are the
 * fused sums of lpm_ic2d_sums with eps = 0 (FP64: the table-driven FP64
 * log; LPM_MIXED: FP32 pairs with FP64 tile carries).  Other arguments as
 * lpm_bve_velocity. */
int lpm_bve_solve(lpm_ctx* ctx, const double* tgt_x, int64_t tgt_es,
                  int64_t tgt_cs, int64_t n_tgt, const double* src_x,
                  int64_t src_es, int64_t src_cs, const double* src_zeta,
                  const double* src_area, const uint8_t* src_mask,
                  int64_t n_src, int collocated, int precision, unsigned flags,
                  double* out_u, int64_t out_es, int64_t out_cs,
                  double* out_psi, void* stream);

/* ---------------------------------------------------------- BVE RK4 ---- */
/* Fields of BVESphere<Seed> the stepper reads and writes (device memory).
 * Positions/velocities use (es, cs) strides as above. */
typedef struct lpm_bve_fields {
  int64_t n_verts;       /* passive particles (mesh vertices) */
  double* vert_x;        /* in/out positions */
  int64_t vert_x_es, vert_x_cs;
  double* vert_zeta;     /* in/out relative vorticity */
  double* vert_u;        /* in: velocity at vert_x; out: at the new vert_x */
  int64_t vert_u_es, vert_u_cs;
  int64_t n_faces;       /* active particles: stored faces incl. divided */
  double* face_x;
  int64_t face_x_es, face_x_cs;
  double* face_zeta;
  double* face_u;
  int64_t face_u_es, face_u_cs;
  const double* face_area;
  const uint8_t* face_mask; /* NULL: no divided faces */
} lpm_bve_fields;

/* Sets vert_u and face_u from the current positions and vorticity
 * (BVESphere::init_velocity, src/lpm_bve_sphere_impl.hpp:181-205). */
int lpm_bve_init_velocity(lpm_ctx* ctx, const lpm_bve_fields* f,
                          int precision, unsigned flags, void* stream);

/* One classical RK4 step of (x, zeta) for vertices and faces, with
 * dzeta/dt = -2 Omega u_z and 4 direct-sum evaluations (BVERK4::
 * advance_timestep, src/lpm_bve_rk4_impl.hpp:55-167; BVERK4Update :12-53;
 * BVEVorticityTendency src/lpm_bve_sphere_kernels.hpp:399-413).
 * On entry vert_u/face_u must hold the velocity at the current positions
 * (from lpm_bve_init_velocity or the previous step); on exit they hold the
 * velocity at the new positions.
 * LPM_RK4_REFERENCE_COMPAT: the face vorticity update uses (k1, k2, k4, k4)
 * as the reference does (src/lpm_bve_rk4_impl.hpp:157); without it,
 * (k1, k2, k3, k4).  LPM_RK4_RENORMALIZE: stage and final positions are
 * projected onto the unit sphere.  LPM_FLAG_SKIP_MASKED_TARGETS: divided
 * faces are neither evaluated nor advanced.  LPM_RK4_USE_GRAPH: the step's
 * launches are captured into a hipGraph on first use and replayed.
 * Stage buffers are owned by the context (sized on first use). */
int lpm_bve_rk4_step(lpm_ctx* ctx, const lpm_bve_fields* f, double dt,
                     double Omega, int precision, unsigned flags,
                     void* stream);

/* -------------------------------------------------------- ensembles ---- */
/* An ensemble of independent BVE problems ("members") stored back to back
 * in one lpm_bve_fields: member m owns vertices [vert_offsets[m],
 * vert_offsets[m + 1]) and faces [face_offsets[m], face_offsets[m + 1]) of
 * the concatenated arrays (strides apply to the concatenated arrays), and
 * its velocity is the sum over its own undivided faces only.  The offset
 * arrays are host memory of n_members + 1 nondecreasing entries from 0 to
 * n_verts / n_faces; every member owns at least one face.
 * All members are evaluated by one persistent kernel that takes
 * (member, target block) work items from a queue, largest members first,
 * so members of very different sizes share one launch.  The work list is
 * uploaded (synchronously) the first time an offset layout is seen and
 * cached in the context.
 * No reference counterpart: the reference evolves one BVESphere at a time
 * (BVESphere::init_velocity, src/lpm_bve_sphere_impl.hpp:181-205;
 * BVERK4::advance_timestep, src/lpm_bve_rk4_impl.hpp:55-167).  Each
 * member's result equals lpm_bve_init_velocity / lpm_bve_rk4_step applied
 * to that member alone, bit for bit. */
typedef struct lpm_bve_ensemble {
  int32_t n_members;
  const int64_t* vert_offsets; /* host, n_members + 1 */
  const int64_t* face_offsets; /* host, n_members + 1 */
  lpm_bve_fields fields;       /* concatenated member arrays */
} lpm_bve_ensemble;

/* Velocity of every member at its current state (the ensemble form of
 * lpm_bve_init_velocity; same flags). */
int lpm_bve_ensemble_init_velocity(lpm_ctx* ctx, const lpm_bve_ensemble* e,
                                   int precision, unsigned flags,
                                   void* stream);

/* One RK4 step of every member (the ensemble form of lpm_bve_rk4_step;
 * same flags, including LPM_RK4_USE_GRAPH, which replays the whole step,
 * with its 4 ensemble launches, from one hipGraph). */
int lpm_bve_ensemble_rk4_step(lpm_ctx* ctx, const lpm_bve_ensemble* e,
                              double dt, double Omega, int precision,
                              unsigned flags, void* stream);

/* ---------------------------------------------- Incompressible2D ---- */
/* Fused velocity and stream function of 2-D incompressible flow on the
 * sphere, with kernel smoothing eps (arg = 1 - x.y + eps^2, Gamma = zeta A):
 *   u(x_i)   = sum_j (x_i cross y_j) Gamma_j / (-4 pi arg),
 *   psi(x_i) = sum_j -log(arg) Gamma_j / (4 pi),
 * over the undivided faces.  Replaces Incompressible2DPassiveSums
 * (collocated = 0) and Incompressible2DActiveSums (collocated = 1) with
 * Incompressible2DReducer and Incompressible2DKernels<SphereGeometry>::
 * kernel_vals (src/lpm_incompressible2d_kernels.hpp:35-47, 96-247).
 * Collocated sums skip j == i only when eps == 0, as the reference does
 * (:242).  precision LPM_FP64, or LPM_MIXED (FP32 pairs on tile-relative
 * offsets with the native FP32 log, FP64 tile carries; no reference
 * counterpart).  Arguments as lpm_bve_velocity; out_u or out_psi may be
 * NULL. */
int lpm_ic2d_sums(lpm_ctx* ctx, const double* tgt_x, int64_t tgt_es,
                  int64_t tgt_cs, int64_t n_tgt, const double* src_x,
                  int64_t src_es, int64_t src_cs, const double* src_zeta,
                  const double* src_area, const uint8_t* src_mask,
                  int64_t n_src, int collocated, double eps, int precision,
                  unsigned flags, double* out_u, int64_t out_es,
                  int64_t out_cs, double* out_psi, void* stream);

/* Fields of Incompressible2D<Seed> (src/lpm_incompressible2d.hpp) the
 * stepper reads and writes. */
typedef struct lpm_ic2d_fields {
  int64_t n_verts;
  double* vert_x;
  int64_t vert_x_es, vert_x_cs;
  double* vert_zeta;
  double* vert_u;
  int64_t vert_u_es, vert_u_cs;
  double* vert_psi;
  int64_t n_faces;
  double* face_x;
  int64_t face_x_es, face_x_cs;
  double* face_zeta;
  double* face_u;
  int64_t face_u_es, face_u_cs;
  double* face_psi;
  const double* face_area;
  const uint8_t* face_mask;
} lpm_ic2d_fields;

/* Velocity and stream function at the current state
 * (Incompressible2D::init_direct_sums,
 * src/lpm_incompressible2d_impl.hpp:235-254). */
int lpm_ic2d_init_direct_sums(lpm_ctx* ctx, const lpm_ic2d_fields* f,
                              double eps, int precision, unsigned flags,
                              void* stream);

/* One step of Incompressible2DRK2 (Heun's method,
 * src/lpm_incompressible2d_rk2_impl.hpp:75-172) with
 * Incompressible2DTendencies (dzeta/dt = -2 Omega u_z, CoriolisSphere;
 * src/lpm_incompressible2d_kernels.hpp:256-277).  On entry u must be
 * current; on exit u and psi are at the new state.  Flags:
 * LPM_FLAG_SKIP_MASKED_TARGETS, LPM_RK4_RENORMALIZE (stage and final
 * positions projected onto the sphere), LPM_RK4_USE_GRAPH. */
int lpm_ic2d_rk2_step(lpm_ctx* ctx, const lpm_ic2d_fields* f, double dt,
                      double Omega, double eps, int precision, unsigned flags,
                      void* stream);

/* --------------------------------------------------------- SWE sums ---- */
/* Sphere shallow-water direct sums at the targets: velocity from relative
 * vorticity zeta and divergence sigma,
 *   u(x) = sum_j kzeta(x, y_j) zeta_j A_j + ksigma(x, y_j) sigma_j A_j,
 * the double dot ddot = grad(u):grad(u)^T of the packed velocity gradient the
 * reference forms, and the Gaussian PSE surface Laplacian
 *   laps(x_i) = eps_p^-4 sum_j [x_i.y_j > 0] (s_j - s_i) A_j (4/pi)
 *               exp(-|x_i - y_j|^2 / eps_p^2).
 * Replaces SphereVertexSums (collocated = 0) / SphereFaceSums (collocated = 1)
 * with SphereSweDirectSumReducer, sphere_swe_velocity_sums, kzeta_sphere,
 * ksigma_sphere, grad_kzeta, grad_ksigma (src/lpm_swe_kernels.hpp:62-362,
 * 578-619, 723-780, 877-930) plus LaplacianReducer / BivariateOrder2
 * (src/lpm_pse.hpp:75-111, 157-203).  The reference's sphere PSE branch is
 * unreachable (its kernels use planar distances); chord distances are used.
 * Collocated sums skip j == i only when eps == 0, as SphereFaceSums does.
 * pse_eps == 0 disables the Laplacian (out_laps untouched).  Any output may
 * be NULL.  With LPM_MIXED, non-collocated targets (and the vertices of the
 * SWE steps and ensembles) are evaluated in the Morton order of their
 * positions, so that each wave's targets lie close together; the order is
 * computed on the host at the first mixed call for a position buffer (that
 * call synchronises the stream once) and reused while the buffer is; a
 * first call made under stream capture keeps the given order. */
int lpm_swe_sphere_sums(lpm_ctx* ctx, const double* tgt_x, int64_t tgt_es,
                        int64_t tgt_cs, int64_t n_tgt, const double* tgt_surf,
                        const double* src_x, int64_t src_es, int64_t src_cs,
                        const double* src_zeta, const double* src_sigma,
                        const double* src_area, const uint8_t* src_mask,
                        const double* src_surf, int64_t n_src, int collocated,
                        double eps, double pse_eps, int precision,
                        unsigned flags, double* out_u, int64_t out_es,
                        int64_t out_cs, double* out_ddot, double* out_laps,
                        void* stream);

/* Fields of SWE<Seed> (src/lpm_swe.hpp:18-177) the sphere stepper reads and
 * writes.  Bottom topography is zero, so surface = depth. */
typedef struct lpm_swe_fields {
  int64_t n_verts;
  double* vert_x;
  int64_t vert_x_es, vert_x_cs;
  double* vert_zeta;   /* state: relative vorticity */
  double* vert_sigma;  /* state: divergence */
  double* vert_h;      /* state: depth */
  double* vert_u;      /* velocity */
  int64_t vert_u_es, vert_u_cs;
  double* vert_ddot;   /* grad(u):grad(u)^T */
  double* vert_laps;   /* surface Laplacian */
  double* vert_surf;   /* surface height */
  int64_t n_faces;
  double* face_x;
  int64_t face_x_es, face_x_cs;
  double* face_zeta;
  double* face_sigma;
  double* face_area;   /* state: area (evolves, dA/dt = sigma A) */
  const double* face_mass; /* fixed mass = depth * area */
  double* face_h;      /* derived depth = mass / area */
  double* face_u;
  int64_t face_u_es, face_u_cs;
  double* face_ddot;
  double* face_laps;
  double* face_surf;
  const uint8_t* face_mask;
} lpm_swe_fields;

/* u, ddot and laps at the current state (SWE::init_direct_sums,
 * src/lpm_swe_impl.hpp:401-445, plus the PSE Laplacian). */
int lpm_swe_init_sums(lpm_ctx* ctx, const lpm_swe_fields* f, double eps,
                      double pse_eps, int precision, unsigned flags,
                      void* stream);

/* One RK4 step of (x, zeta, sigma, h | A) in the stage layout of
 * SWERK4::advance_timestep (src/lpm_swe_rk4_impl.hpp:204-445; update
 * SWERK4Update :12-87; tendencies SWEVorticityDivergenceHeightTendencies /
 * ...AreaTendencies src/lpm_swe_kernels.hpp:941-1069 with CoriolisSphere
 * src/lpm_coriolis.hpp:154-195), with the sphere sums.  On entry u, ddot,
 * laps and surf must be current (lpm_swe_init_sums or the previous step).
 * LPM_RK4_REFERENCE_COMPAT: the stage-4 position increment stays zero, as
 * in the reference, which never assigns it.  LPM_FLAG_SKIP_MASKED_TARGETS:
 * divided faces are neither evaluated nor advanced.  LPM_RK4_USE_GRAPH:
 * the step is captured into a hipGraph on first use and replayed (needs a
 * non-default stream). */
int lpm_swe_rk4_step(lpm_ctx* ctx, const lpm_swe_fields* f, double dt,
                     double Omega, double g, double eps, double pse_eps,
                     int precision, unsigned flags, void* stream);

/* One RK2 (Heun) step in the stage layout of SWERK2::advance_timestep_impl
 * (src/lpm_swe_rk2_impl.hpp:80-257): increments at the state, the work
 * state = state + increments with its surface (SetSurfaceFromDepth /
 * SetDepthAndSurfaceFromMassAndArea), sums and increments there, then
 * state += increments1/2 + increments2/2 and the sums at the new state.
 * The surface Laplacian comes from the PSE sums (pse_eps), where the
 * reference evaluates a GMLS operator (Compadre).  Preconditions and flags
 * as lpm_swe_rk4_step (LPM_RK4_REFERENCE_COMPAT has no effect). */
int lpm_swe_rk2_step(lpm_ctx* ctx, const lpm_swe_fields* f, double dt,
                     double Omega, double g, double eps, double pse_eps,
                     int precision, unsigned flags, void* stream);

/* Ensembles of sphere SWE problems stored back to back, as
 * lpm_bve_ensemble (same offset rules): one persistent launch evaluates the
 * sums of every member.  No reference counterpart; each member's results
 * equal lpm_swe_init_sums / lpm_swe_rk4_step on that member alone, bit for
 * bit. */
typedef struct lpm_swe_ensemble {
  int32_t n_members;
  const int64_t* vert_offsets; /* host, n_members + 1 */
  const int64_t* face_offsets; /* host, n_members + 1 */
  lpm_swe_fields fields;       /* concatenated member arrays */
} lpm_swe_ensemble;

int lpm_swe_ensemble_init_sums(lpm_ctx* ctx, const lpm_swe_ensemble* e,
                               double eps, double pse_eps, int precision,
                               unsigned flags, void* stream);
int lpm_swe_ensemble_rk4_step(lpm_ctx* ctx, const lpm_swe_ensemble* e,
                              double dt, double Omega, double g, double eps,
                              double pse_eps, int precision, unsigned flags,
                              void* stream);
int lpm_swe_ensemble_rk2_step(lpm_ctx* ctx, const lpm_swe_ensemble* e,
                              double dt, double Omega, double g, double eps,
                              double pse_eps, int precision, unsigned flags,
                              void* stream);

/* ---------------------------------------------------- kernel variants ---- */
/* The direct-sum kernels are compiled in several tile/block/unroll shapes
 * ("variants").  No reference counterpart (Kokkos picks TeamPolicy(n, AUTO),
 * src/lpm_bve_rk4_impl.hpp:70-71).  kernel_id is LPM_KERNEL_*.
 * BVE velocity variants named *_splitK sum each target block's sources in
 * K chunks and add the partial sums in a second kernel (same results to
 * summation order); they keep a workspace of 3 K doubles per target in the
 * context.  lpm_bve_init_velocity and lpm_bve_rk4_step run them (the step
 * sizes the workspace before any graph capture); lpm_bve_velocity and
 * lpm_bve_velocity_noalloc run the unsplit variant of the same shape. */
int lpm_variant_count(int kernel_id);
const char* lpm_variant_name(int kernel_id, int variant);
/* The fixed variant used when nothing was tuned or forced. */
int lpm_variant_default(int kernel_id);
/* Forces `variant` for every problem size; -1 returns to the tuned/default
 * choice. */
int lpm_set_variant(lpm_ctx* ctx, int kernel_id, int variant);
/* Autotuner: times every variant on the velocity evaluation of `f` (it
 * overwrites vert_u/face_u with the velocity at the current state, as
 * lpm_bve_init_velocity does), `repeats` launches each, and records the
 * fastest for this device and problem-size bucket (floor(log2(targets))).
 * all_ms_host (nullable) receives lpm_variant_count() per-variant times. */
int lpm_bve_autotune(lpm_ctx* ctx, const lpm_bve_fields* f, int precision,
                     unsigned flags, int repeats, int* best_variant,
                     double* best_ms, double* all_ms_host, void* stream);
/* The same for the ensemble kernel (LPM_KERNEL_BVE_ENSEMBLE_*), keyed by
 * the ensemble's total target count. */
int lpm_bve_ensemble_autotune(lpm_ctx* ctx, const lpm_bve_ensemble* e,
                              int precision, unsigned flags, int repeats,
                              int* best_variant, double* best_ms,
                              double* all_ms_host, void* stream);
/* Tuning table persistence (text: "arch/kernel/bucket variant" lines). */
int lpm_tune_save(lpm_ctx* ctx, const char* path);
int lpm_tune_load(lpm_ctx* ctx, const char* path);

/* ------------------------------------------------------- diagnostics ---- */
/* Relative error norms of a device field against its exact values:
 * replaces the ErrNorms constructors with their error computation and
 * parallel_reduce (src/lpm_error.hpp:81-131, src/lpm_error_impl.hpp:61-182).
 * Element i of appx / exact / err has ncomp (1..3) components at
 * base[i * es + k * cs]; e_i = (mask && mask[i]) ? 0 : appx_i - exact_i,
 * written to err when err is not NULL (err may be appx).  weight (NULL: 1)
 * weighs the l1 and l2 sums; masked elements keep their exact magnitude in
 * the denominators.  norms (host) receives {l1, l2, linf}, each quotient
 * num * den / (den^2 + eps^2) with eps the FP64 machine epsilon.  Blocking
 * on the stream, as the reference's reduction to the host is. */
int lpm_err_norms(lpm_ctx* ctx, const double* appx, int64_t appx_es,
                  int64_t appx_cs, const double* exact, int64_t exact_es,
                  int64_t exact_cs, const double* weight, const uint8_t* mask,
                  int64_t n, int ncomp, double* err, int64_t err_es,
                  int64_t err_cs, double* norms, void* stream);
/* u_i . x_i per particle into out (device, n values): SphereTangentFunctor
 * (src/lpm_sphere_functions.hpp:62-78).  Stream-ordered. */
int lpm_sphere_tangency(lpm_ctx* ctx, const double* u, int64_t u_es,
                        int64_t u_cs, const double* x, int64_t x_es,
                        int64_t x_cs, int64_t n, double* out, void* stream);

/* ------------------------------------------------------------ timing ---- */
/* HIP events recorded around every direct-sum launch, on the stream the
 * kernel is launched on (disabled inside graph capture).  Counterpart of the
 * reference's Timer with Kokkos::fence (src/util/lpm_timer.hpp:16-51).
 * lpm_timing_enable resets all records; lpm_timing_query synchronises on the
 * pending events and returns the accumulated kernel time, launch count and
 * pair count (targets x stored sources) of one kernel family. */
#define LPM_KERNEL_BVE_VEL_F64 0
#define LPM_KERNEL_BVE_VEL_MIXED 1
#define LPM_KERNEL_BVE_PSI 2
#define LPM_KERNEL_SWE_F64 3
#define LPM_KERNEL_SWE_MIXED 4
#define LPM_KERNEL_BVE_ENSEMBLE_F64 5
#define LPM_KERNEL_BVE_ENSEMBLE_MIXED 6
#define LPM_KERNEL_IC2D 7
int lpm_timing_enable(lpm_ctx* ctx, int enable);
int lpm_timing_query(lpm_ctx* ctx, int kernel_id, double* total_ms,
                     int64_t* launches, double* pairs);

#ifdef __cplusplus
}
#endif

#endif /* LPM_MI355X_H */
