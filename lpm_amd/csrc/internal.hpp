// Internal declarations shared by the HIP translation units of
// liblpm_mi355x.so.  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lpm_mi355x.h"

namespace lpm_amd {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define LPM_HIP_CHECK(expr)                                                 \
  do {                                                                      \
    hipError_t lpm_e_ = (expr);                                             \
    if (lpm_e_ != hipSuccess)                                               \
      throw ::lpm_amd::Error(LPM_ERR_HIP, std::string(#expr) + " failed: " + \
                                              hipGetErrorString(lpm_e_));   \
  } while (0)

#define LPM_ARG_CHECK(cond, msg)                                  \
  do {                                                            \
    if (!(cond)) throw ::lpm_amd::Error(LPM_ERR_ARG, (msg));      \
  } while (0)

// Incremented whenever any workspace is (re)allocated; captured graphs
// record it and are re-captured when it changed, so that a replay never
// touches freed workspace memory.
extern std::atomic<uint64_t> g_workspace_epoch;

// Grow-only device buffer owned by a context.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  void* ensure(size_t nbytes);  // reallocates (contents lost) when too small
  template <typename T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

// Strided coordinate/vector arrays: element (i, k) at base[i*es + k*cs].
struct CrdIn {
  const double* base;
  int64_t es, cs;
};
struct CrdOut {
  double* base;
  int64_t es, cs;
};

// Elements per block of the mask scan.
constexpr int kMaskScanChunk = 1024;

// Compaction of a face mask: rank[i] = packed index of face i (-1 if
// divided), inv[r] = face of packed index r, *d_count = number packed,
// block_offsets[b] = number of undivided faces before element
// b * kMaskScanChunk (b = 0 .. ceil(n / kMaskScanChunk)).
// With no mask, rank and inv are the identity and not materialised.
struct RankMap {
  const int32_t* rank = nullptr;
  const int32_t* inv = nullptr;
  const int32_t* d_count = nullptr;
  const int32_t* block_offsets = nullptr;
  int64_t n = 0;
};

// A target set of one direct-sum launch.  Thread slot c (0 <= c < n, or
// < *d_n when d_n is set) evaluates target i = index ? index[c] : c.
struct TargetSeg {
  CrdIn x{nullptr, 0, 0};
  int64_t n = 0;
  const int32_t* index = nullptr;      // slot -> target (compacted targets)
  const int32_t* d_n = nullptr;        // device count of live slots
  const int32_t* self_rank = nullptr;  // collocated: packed index of target
                                       // i's own source, -1 if none
  bool self_identity = false;          // collocated: self = slot c
  const uint8_t* skip = nullptr;       // targets to leave untouched
  CrdOut out{nullptr, 0, 0};           // velocity output
  double* out_scalar = nullptr;        // stream-function output
};

// Sources after the prologue: undivided faces, compacted and packed.
struct PackedSources {
  const void* data = nullptr;      // double4 {x,y,z,s} (FP64) or float4
                                   // {x',y',z',2s} tile-relative (mixed)
  const double4* centers = nullptr; // mixed: tile centres (x, y, z, -)
  const double4* sy = nullptr;     // FP64: {s y, 0} per source
  const int32_t* count = nullptr;  // device: number of packed sources
  int64_t capacity = 0;            // host upper bound (n stored faces)
  // mixed: matrix-core layout (see bve_device.hpp)
  const float4* mf_y = nullptr;
  const float* mf_q = nullptr;
  const float* mf_s = nullptr;
  const float* mf_t = nullptr;
  const double4* mf_c = nullptr;
  // mixed: f16 hi/lo distance operands and the 32x32 weight operand
  const uint4* mf_a = nullptr;
  const float* mf_t32 = nullptr;
  const double* mf_is = nullptr;
};

// Kernel timing records (HIP events around direct-sum launches).
enum KernelId : int {
  KID_BVE_VEL_F64 = 0,
  KID_BVE_VEL_MIXED = 1,
  KID_BVE_PSI = 2,
  KID_SWE_F64 = 3,
  KID_SWE_MIXED = 4,
  KID_BVE_ENS_F64 = 5,
  KID_BVE_ENS_MIXED = 6,
  KID_IC2D = 7,
  KID_COUNT = 8
};

struct TimingRec {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double total_ms = 0;
  int64_t launches = 0;
  double pairs = 0;
};

// Launch configuration of the tiled direct-sum kernels.
struct VelConfig {
  int variant = 0;  // index into the compiled kernel table
};

}  // namespace lpm_amd

struct lpm_ctx {
  int device = 0;
  std::string last_error;
  int num_cus = 0;
  std::string arch;

  // prologue workspaces
  lpm_amd::DevBuf rank;         // int32 per stored face
  lpm_amd::DevBuf inv;          // int32 per stored face (packed -> face)
  lpm_amd::DevBuf block_counts; // int32 per scan block (+1 total)
  lpm_amd::DevBuf packed64;     // double4 per source
  lpm_amd::DevBuf packed64sy;   // double4 {s y, 0} per source (FP64)
  lpm_amd::DevBuf split_part;   // chunk partials of source-split sums
  lpm_amd::DevBuf packed32;     // float4 per source (mixed)
  lpm_amd::DevBuf centers;      // double4 per tile (mixed)
  lpm_amd::DevBuf mf_y, mf_q, mf_s, mf_t, mf_c;  // matrix-core layout (mixed)
  lpm_amd::DevBuf mf_a, mf_t32, mf_is;          // f16-split layout (mixed)
  lpm_amd::DevBuf flag;         // int32 non-finite flag
  lpm_amd::DevBuf diag;         // error-norm partial sums
  lpm_amd::DevBuf count_ident;  // int32 count when no mask

  lpm_amd::DevBuf swe_packed;   // SWE sources (64 B each)
  lpm_amd::DevBuf swe_bounds;   // SWE tile bounding spheres
  lpm_amd::DevBuf swe_feat;     // SWE matrix-core tile records
  lpm_amd::DevBuf swe_far;      // SWE two-pass far sums (18 per target)

  // ensemble work plans (keyed by targets per work item) and scratch
  struct EnsemblePlan;
  std::map<int, std::shared_ptr<EnsemblePlan>> ens_plans;
  uint64_t ens_plan_serial = 0;
  lpm_amd::DevBuf ens_bounds;   // int32 per member boundary
  lpm_amd::DevBuf ens_queue;    // int32 work-queue head, padded total

  // RK4 stage buffers
  lpm_amd::DevBuf rk_v, rk_f;   // BVE vertex / face stage storage
  lpm_amd::DevBuf swe_v, swe_f; // SWE vertex / face stage storage
  lpm_amd::DevBuf ic2d_v, ic2d_f; // Incompressible2D RK2 stage storage

  // timing
  bool timing = false;
  lpm_amd::TimingRec timing_rec[lpm_amd::KID_COUNT];
  std::vector<hipEvent_t> event_pool;

  // graph cache for RK4 steps
  struct GraphEntry;
  std::map<std::string, std::shared_ptr<GraphEntry>> graphs;
  uint64_t graph_clock = 0;  // LRU stamps of the graph cache

  // stream of the last call that used the context's workspaces, and the
  // event that orders a call on another stream after it
  hipStream_t last_stream = nullptr;
  bool has_last_stream = false;
  hipEvent_t order_event = nullptr;

  // autotuned kernel variants, keyed by (arch, kernel, size bucket)
  std::map<std::string, int> tuned;
  int forced[lpm_amd::KID_COUNT] = {-1, -1, -1, -1, -1, -1, -1, -1};
  // bumped whenever `tuned` or `forced` changes: captured graphs record the
  // kernels of the choice in force when they were captured
  uint64_t choice_serial = 0;

  // spatially compact slot orders of target sets (target_order)
  struct TargetOrder {
    const double* x = nullptr;
    int64_t es = 0, cs = 0, n = 0;
    uint64_t ranges = 0;  // hash of the range offsets
    uint64_t last_use = 0;
    lpm_amd::DevBuf perm;  // int32 per slot: target index
  };
  std::vector<std::unique_ptr<TargetOrder>> orders;
  uint64_t order_clock = 0;
};

namespace lpm_amd {

void set_error(lpm_ctx* ctx, const std::string& msg);

// Orders this call's work on stream s after everything the context issued
// on the stream of its previous call (the workspaces are shared by all of
// a context's calls).  No-op when the stream is unchanged or capturing.
void order_stream(lpm_ctx* ctx, hipStream_t s);

// Kernel variants of a family (KID_*) and the autotuning table.
int variant_count(int family);
const char* variant_name(int family, int v);
int default_variant(int family);
std::string tune_key(const lpm_ctx* ctx, int family, int64_t n_targets);
int pick_variant(lpm_ctx* ctx, int family, int64_t n_targets);
hipStream_t as_stream(void* s);

// Event bracket for a timed launch (no-op when timing is off).
struct TimedLaunch {
  lpm_ctx* ctx;
  int kid;
  hipStream_t stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double pairs;
  TimedLaunch(lpm_ctx* c, int k, hipStream_t s, double p);
  ~TimedLaunch() noexcept(false);
};

// ---- prologue (bve_sums.hip) ----
RankMap build_rank(lpm_ctx* ctx, const uint8_t* mask, int64_t n,
                   hipStream_t s);
PackedSources pack_sources(lpm_ctx* ctx, CrdIn sx, const double* zeta,
                           const double* area, const RankMap& map,
                           int precision, hipStream_t s);

// The same prologue into caller-provided device buffers (no allocation).
// Sizes for n stored faces: counts ceil(n / kMaskScanChunk) + 1, rank and
// inv n each (unused without a mask), count_ident 1.
struct RankWs {
  int32_t* counts = nullptr;
  int32_t* rank = nullptr;
  int32_t* inv = nullptr;
  int32_t* count_ident = nullptr;
};
RankMap build_rank_into(const RankWs& ws, const uint8_t* mask, int64_t n,
                        hipStream_t s);

// Source layouts of the velocity variants (bit set).  Every layout also
// needs the FP64 packing p64.  With tcap = 256 * ceil(n / 256) and
// ntile = tcap / 256 the arrays hold: p64, psy n double4; p32 n float4;
// centers, mf_c ntile double4; mf_y tcap float4; mf_q, mf_s tcap float;
// mf_t, mf_t32 4 tcap float; mf_a 2 tcap uint4; mf_is ntile double.
enum : unsigned {
  kLayoutSy = 1u,      // FP64 {s y} (psy)
  kLayoutVec32 = 2u,   // mixed vector-ALU kernels (p32, centers)
  kLayoutMfma = 4u,    // f32 matrix-core kernel (mf_y, mf_q, mf_s, mf_t, mf_c)
  kLayoutMfma16 = 8u,  // f16-split kernels (mf_y, mf_s, mf_c, mf_a, mf_t32,
                       // mf_is)
};
struct PackWs {
  double4* p64 = nullptr;
  double4* psy = nullptr;
  float4* p32 = nullptr;
  double4* centers = nullptr;
  float4* mf_y = nullptr;
  float* mf_q = nullptr;
  float* mf_s = nullptr;
  float* mf_t = nullptr;
  double4* mf_c = nullptr;
  uint4* mf_a = nullptr;
  float* mf_t32 = nullptr;
  double* mf_is = nullptr;
};
PackedSources pack_sources_into(const PackWs& ws, unsigned layouts, CrdIn sx,
                                const double* zeta, const double* area,
                                const RankMap& map, int precision,
                                hipStream_t s);
// Variant of the BVE velocity kernel a launch over n_targets would use, and
// the source layouts it reads.
int bve_velocity_variant(lpm_ctx* ctx, int precision, int64_t n_targets);
unsigned bve_variant_layouts(int precision, int variant);
// Face targets of a collocated sum: every stored face (self skipped by
// rank) or, with skip_masked, only the undivided ones (compacted slots).
TargetSeg face_targets(CrdIn x, const RankMap& map, const uint8_t* mask,
                       bool skip_masked, CrdOut out, double* out_scalar);
// variant < 0: bve_velocity_variant's choice.  timed = false skips the
// timing events.  allow_split = false runs a source-split variant as its
// unsplit twin (same shape, one summation level less, no workspace).
void launch_bve_velocity(lpm_ctx* ctx, const PackedSources& src,
                         TargetSeg* segs, int nseg, double eps2, int precision,
                         hipStream_t s, int variant = -1, bool timed = true,
                         bool allow_split = false);
// Bytes of partial sums a source-split variant needs for n_targets (0 for
// every other variant).
size_t bve_split_bytes(int precision, int variant, int64_t n_targets);
// The context's split workspace grown to `need` bytes, or nullptr when it
// cannot be: s is capturing and it is smaller, or the device is out of
// memory (the caller then runs an unsplit variant).
double* split_workspace(lpm_ctx* ctx, size_t need, hipStream_t s);
// Source chunks per target block for a launch of `blocks` target blocks of
// `kernel` over nsrc_upper sources in tiles of `tile`: enough that the
// items fill eight rounds of the blocks the device holds at once, with at
// least 8 tiles per chunk (1: no split).
int split_chunks(lpm_ctx* ctx, const void* kernel, int block, int64_t blocks,
                 int64_t nsrc_upper, int tile);
void launch_bve_stream_fn(lpm_ctx* ctx, const PackedSources& src,
                          TargetSeg* segs, int nseg, hipStream_t s);
void check_finite(lpm_ctx* ctx, const double* base, int64_t n, int64_t es,
                  int64_t cs, int ncomp, hipStream_t s);
// check_finite with a caller-provided device flag (one int32).
void check_finite_into(int32_t* flag, const double* base, int64_t n,
                       int64_t es, int64_t cs, int ncomp, hipStream_t s);

// ---- allocation-free velocity (noalloc.hip) ----
// Looks up the device's static scratch array (once per device; the device
// must be current).  Failures are recorded and reported by the entry point.
void noalloc_setup(int device);

// ---- RK4 (bve_rk4.hip) ----
void validate_bve_fields(const lpm_bve_fields* f);
// Runs issue() on stream s, or with use_graph replays the hipGraph captured
// from it under key_fn() (which must also allocate every workspace issue()
// uses; the graph is re-captured when a workspace was reallocated).
void run_maybe_graph(lpm_ctx* ctx, bool use_graph, hipStream_t s,
                     const std::function<std::string()>& key_fn,
                     const std::function<void()>& issue);

// ---- ensembles (ensemble.hip, bve_ensemble.hip) ----
// Host view of an ensemble's member offsets (validated).
struct EnsembleDesc {
  int n_members = 0;
  const int64_t* voff = nullptr;  // host, n_members + 1
  const int64_t* foff = nullptr;  // host, n_members + 1
  int64_t n_verts = 0, n_faces = 0;
  int64_t max_member_faces = 0;
  int64_t total_padded_upper = 0;  // sum of member faces rounded up to 256
};
EnsembleDesc validate_members(int32_t n_members, const int64_t* voff,
                              const int64_t* foff, int64_t n_verts,
                              int64_t n_faces);

// Device member table entry.  v0/nv/f0/nf come from the host offsets; R
// (first compacted rank of the member's undivided faces), cnt (their
// number) and P (first packed slot, a multiple of kEnsTile) are computed on
// the device from the face mask at every evaluation.
constexpr int kEnsTile = 256;
struct EnsMember {
  int64_t v0, nv, f0, nf;
  int32_t R, cnt, P, pad;
};

// Work plan of one offset layout for `per` targets per work item: items
// {member, segment (0 vertices, 1 faces), block, 0} ordered by decreasing
// member face count, and the member table.
struct EnsLayout {
  EnsMember* mem = nullptr;
  const int4* items = nullptr;
  int n_items = 0;
  int32_t* queue = nullptr;         // work-queue head
  int32_t* padded_total = nullptr;  // packed slots in use (device)
  double pairs = 0;                 // sum of (member targets) x (member faces)
  uint64_t serial = 0;              // changes whenever the plan is rebuilt
};
// Plan and workspaces only (no launches): safe before a stream capture.
EnsLayout ensemble_plan(lpm_ctx* ctx, const EnsembleDesc& d, int per);
// ensemble_plan plus the device member table from the mask (launches on s).
EnsLayout ensemble_layout(lpm_ctx* ctx, const EnsembleDesc& d,
                          const RankMap& map, const uint8_t* face_mask,
                          int per, hipStream_t s);
void ensemble_reset_queue(const EnsLayout& L, hipStream_t s);
// Persistent grid: resident blocks of `kernel` on the device, <= n_items.
int ensemble_grid(lpm_ctx* ctx, const void* kernel, int block, int n_items);

int ens_variant_count(int family);
const char* ens_variant_name(int family, int v);
int ens_default_variant(int family);
// BVE: allocates every workspace an evaluation uses (so that a stream
// capture that follows allocates nothing); returns the plan serial.
uint64_t ensemble_prepare(lpm_ctx* ctx, const EnsembleDesc& d,
                          const lpm_bve_fields& f, int precision,
                          hipStream_t s);
// BVE velocity of every member from positions (vx, fx) and face vorticity
// fzeta, written to f.vert_u / f.face_u.
void ensemble_velocity(lpm_ctx* ctx, const EnsembleDesc& d,
                       const lpm_bve_fields& f, CrdIn vx, CrdIn fx,
                       const double* fzeta, const RankMap& map, int precision,
                       unsigned flags, hipStream_t s);

// ---- sphere SWE sums (swe_sums.hip) ----
// Kernel variants of KID_SWE_F64 / KID_SWE_MIXED (the default is 0).
int swe_variant_count(int family);
const char* swe_variant_name(int family, int v);
struct SwePacked {
  const void* data = nullptr;       // 64 B per source
  const double4* bounds = nullptr;  // per 256-source tile: centre, radius^2
  const int32_t* count = nullptr;
  int64_t capacity = 0;
};
struct SweTargets {
  TargetSeg base;                   // positions, slots, self, velocity out
  const double* surf = nullptr;     // target surface height (PSE)
  double* ddot = nullptr;
  double* laps = nullptr;
};
// Allocates the matrix-core workspaces a mixed launch over n_src stored
// faces and n_targets targets may use (so that a stream capture that follows
// allocates nothing).
void swe_feature_workspace(lpm_ctx* ctx, int64_t n_src, int64_t n_targets,
                           int precision);
SwePacked swe_pack_sources(lpm_ctx* ctx, CrdIn sx, const double* zeta,
                           const double* sigma, const double* area,
                           const double* surf, const RankMap& map,
                           hipStream_t s);
void launch_swe_sums(lpm_ctx* ctx, const SwePacked& src, SweTargets* segs,
                     int nseg, double eps, double pse_eps, int precision,
                     hipStream_t s);
// Slot order of n targets at positions x (target indices, int32 on the
// device) that keeps each wave's targets close together: Morton order of
// the positions within each range [off[m], off[m + 1]) (one range when off
// is null).  Computed on the host from the positions at the first call for
// these buffers (it synchronises the stream) and kept in the context (at
// most kOrderCacheMax orders); later calls reuse it as the positions move,
// which changes only how targets are grouped.  Returns nullptr (the given
// order) for n == 0 and for a first call made while s is being captured.
constexpr int kOrderCacheMax = 8;
const int32_t* target_order(lpm_ctx* ctx, CrdIn x, int64_t n,
                            const int64_t* off, int n_ranges, hipStream_t s);

// SWE ensembles: workspaces only (returns the plan serial), and the sums of
// every member at positions (vx, fx) from its own undivided faces.
uint64_t swe_ensemble_prepare(lpm_ctx* ctx, const EnsembleDesc& d,
                              const lpm_swe_fields& f, int precision,
                              hipStream_t s);
void swe_ensemble_sums(lpm_ctx* ctx, const EnsembleDesc& d,
                       const lpm_swe_fields& f, CrdIn vx, CrdIn fx,
                       const double* fzeta, const double* fsigma,
                       const double* farea, const RankMap& map, double eps,
                       double pse_eps, int precision, unsigned flags,
                       hipStream_t s);

}  // namespace lpm_amd
