// Allocation-free BVE velocity: lpm_bve_velocity with every workspace carved
// from one file-scope device array instead of context buffers, for callers
// that may not call hipMalloc (stream capture, serving loops).
//
// The array lives in the code object: each device that loads the library
// holds its own copy (kNoAllocScratchBytes) until the process ends.  Host
// code never uses the array's name as an address; noalloc_setup() looks up
// each device's copy with hipGetSymbolAddress once, with that device
// current, and the entry point passes that pointer to the kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "bve_device.hpp"
#include "internal.hpp"

namespace lpm_amd {

namespace {

// Largest source set (stored faces, divided ones included) one call takes.
constexpr int64_t kNoAllocMaxSources = int64_t(1) << 19;
constexpr int64_t kAlign = 256;
constexpr int kMaxDevices = 64;

constexpr int64_t align_up(int64_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// Byte offsets of every workspace of one call in the scratch array; a
// negative offset marks a buffer the call does not use.
struct Carve {
  int64_t counts = -1, rank = -1, inv = -1, count_ident = -1, flag = -1;
  int64_t p64 = -1, psy = -1, p32 = -1, centers = -1;
  int64_t mf_y = -1, mf_q = -1, mf_s = -1, mf_t = -1, mf_c = -1;
  int64_t mf_a = -1, mf_t32 = -1, mf_is = -1;
  int64_t total = 0;
};

// Sizes as the prologue documents them: counts ceil(n / kMaskScanChunk) + 1
// int32, rank and inv n int32 (with a mask); p64, psy n double4; p32 n
// float4; centers, mf_c ntile double4; mf_y tcap float4; mf_q, mf_s tcap
// float; mf_t, mf_t32 4 tcap float; mf_a 2 tcap uint4; mf_is ntile double,
// with tcap = 256 ceil(n / 256) and ntile = tcap / 256.
constexpr Carve carve_scratch(int64_t n, bool mask, unsigned layouts,
                              bool fp64) {
  Carve c{};
  const int64_t cap = n > 1 ? n : 1;
  const int64_t ntile = (cap + kPackTile - 1) / kPackTile;
  const int64_t tcap = ntile * kPackTile;
  int64_t off = 0;
  auto take = [&off](int64_t bytes) {
    const int64_t o = off;
    off += align_up(bytes);
    return o;
  };
  c.flag = take(4);
  if (mask) {
    c.counts = take(4 * ((cap + kMaskScanChunk - 1) / kMaskScanChunk + 1));
    c.rank = take(4 * cap);
    c.inv = take(4 * cap);
  } else {
    c.count_ident = take(4);
  }
  c.p64 = take(32 * cap);
  if (fp64) {
    if (layouts & kLayoutSy) c.psy = take(32 * cap);
  } else {
    if (layouts & kLayoutVec32) {
      c.p32 = take(16 * cap);
      c.centers = take(32 * ntile);
    }
    if (layouts & (kLayoutMfma | kLayoutMfma16)) {
      c.mf_y = take(16 * tcap);
      c.mf_s = take(4 * tcap);
      c.mf_c = take(32 * ntile);
    }
    if (layouts & kLayoutMfma) {
      c.mf_q = take(4 * tcap);
      c.mf_t = take(16 * tcap);
    }
    if (layouts & kLayoutMfma16) {
      c.mf_a = take(32 * tcap);
      c.mf_t32 = take(16 * tcap);
      c.mf_is = take(8 * ntile);
    }
  }
  c.total = off;
  return c;
}

constexpr int64_t cmax(int64_t a, int64_t b) { return a > b ? a : b; }

// Worst case over the layouts of every kernel variant at the maximum size.
constexpr int64_t kNoAllocScratchBytes = cmax(
    cmax(carve_scratch(kNoAllocMaxSources, true, kLayoutSy, true).total,
         carve_scratch(kNoAllocMaxSources, true, kLayoutVec32, false).total),
    cmax(carve_scratch(kNoAllocMaxSources, true, kLayoutMfma, false).total,
         carve_scratch(kNoAllocMaxSources, true, kLayoutMfma16, false).total));
static_assert(kNoAllocScratchBytes <= (int64_t(56) << 20),
              "the allocation-free scratch array stays under 56 MiB");

}  // namespace

}  // namespace lpm_amd

// One copy per device, zero-initialised at load.
static __device__ __align__(256) unsigned char
    scratch[lpm_amd::kNoAllocScratchBytes];

namespace lpm_amd {

namespace {

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};
std::string g_scratch_error[kMaxDevices];

template <typename T>
T* at(void* base, int64_t off) {
  return off < 0 ? nullptr
                 : reinterpret_cast<T*>(static_cast<unsigned char*>(base) + off);
}

void* scratch_ptr(int device) {
  LPM_ARG_CHECK(device >= 0 && device < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device])
    throw Error(LPM_ERR_STATE,
                "allocation-free scratch of device " + std::to_string(device) +
                    " is not set up: " +
                    (g_scratch_error[device].empty()
                         ? std::string("no context was created on it")
                         : g_scratch_error[device]));
  return g_scratch_ptr[device];
}

void check_crd(const double* p, int64_t es, int64_t cs, int64_t n,
               const char* what) {
  if (n == 0) return;
  LPM_ARG_CHECK(p != nullptr, std::string(what) + " is NULL");
  LPM_ARG_CHECK(es > 0 && cs > 0, std::string(what) + ": strides must be > 0");
}

}  // namespace

void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_scratch_error[device] =
        std::string("hipGetSymbolAddress failed: ") + hipGetErrorString(e);
    return;
  }
  g_scratch_ptr[device] = p;
  g_scratch_error[device].clear();
}

}  // namespace lpm_amd

using namespace lpm_amd;

extern "C" {

int64_t lpm_bve_velocity_noalloc_max_sources(void) {
  return kNoAllocMaxSources;
}

int lpm_bve_velocity_noalloc(lpm_ctx* ctx, const double* tgt_x, int64_t tgt_es,
                             int64_t tgt_cs, int64_t n_tgt,
                             const double* src_x, int64_t src_es,
                             int64_t src_cs, const double* src_zeta,
                             const double* src_area, const uint8_t* src_mask,
                             int64_t n_src, int collocated, double eps,
                             int precision, unsigned flags, double* out_u,
                             int64_t out_es, int64_t out_cs, void* stream) {
  try {
    LPM_ARG_CHECK(ctx, "null context");
    LPM_ARG_CHECK(n_tgt >= 0 && n_src >= 0, "negative size");
    LPM_ARG_CHECK(n_tgt < (int64_t(1) << 31), "target count out of range");
    LPM_ARG_CHECK(n_src <= kNoAllocMaxSources,
                  "n_src = " + std::to_string(n_src) +
                      " exceeds the allocation-free maximum of " +
                      std::to_string(kNoAllocMaxSources) +
                      " stored faces; use lpm_bve_velocity");
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
    void* base = scratch_ptr(ctx->device);
    hipStream_t s = as_stream(stream);

    // the variant lpm_bve_velocity would run, and only the layouts it reads
    const int variant = bve_velocity_variant(ctx, precision, n_tgt);
    const unsigned layouts = bve_variant_layouts(precision, variant);
    const Carve c =
        carve_scratch(n_src, src_mask != nullptr, layouts, precision == LPM_FP64);
    LPM_ARG_CHECK(c.total <= kNoAllocScratchBytes,
                  "allocation-free scratch too small for this call");

    RankWs rw;
    rw.counts = at<int32_t>(base, c.counts);
    rw.rank = at<int32_t>(base, c.rank);
    rw.inv = at<int32_t>(base, c.inv);
    rw.count_ident = at<int32_t>(base, c.count_ident);
    const RankMap map = build_rank_into(rw, src_mask, n_src, s);

    PackWs pw;
    pw.p64 = at<double4>(base, c.p64);
    pw.psy = at<double4>(base, c.psy);
    pw.p32 = at<float4>(base, c.p32);
    pw.centers = at<double4>(base, c.centers);
    pw.mf_y = at<float4>(base, c.mf_y);
    pw.mf_q = at<float>(base, c.mf_q);
    pw.mf_s = at<float>(base, c.mf_s);
    pw.mf_t = at<float>(base, c.mf_t);
    pw.mf_c = at<double4>(base, c.mf_c);
    pw.mf_a = at<uint4>(base, c.mf_a);
    pw.mf_t32 = at<float>(base, c.mf_t32);
    pw.mf_is = at<double>(base, c.mf_is);
    const PackedSources ps =
        pack_sources_into(pw, layouts, CrdIn{src_x, src_es, src_cs}, src_zeta,
                          src_area, map, precision, s);

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
    // untimed: the timing records create events
    launch_bve_velocity(ctx, ps, &t, 1, eps * eps, precision, s, variant,
                        false);
    if (flags & LPM_FLAG_CHECK_FINITE)
      check_finite_into(at<int32_t>(base, c.flag), out_u, n_tgt, out_es,
                        out_cs, 3, s);
    return LPM_OK;
  } catch (const Error& e) {
    set_error(ctx, e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_error(ctx, e.what());
    return LPM_ERR_INTERNAL;
  }
}

}  // extern "C"
