// Host-side internals of the HIP engine shared by its translation units
// (poa_hip.hip: the per-call and batched entry points; resident.hip: the
// device-resident window form): status strings, the score-width choice,
// launch configurations, device and pinned buffers, and the engine itself.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "poa_types.hpp"
#include "spoa_hip.h"
#include "thread_pool.hpp"

namespace spoa_hip_host {

using spoa_hip_dev::DevJob;
using spoa_hip_dev::KArgs;
using spoa_hip_dev::KParams;


inline thread_local std::string g_last_error;

inline int Fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

inline bool HipOk(hipError_t rc, const char* what) {
  if (rc == hipSuccess) return true;
  g_last_error = std::string("[spoa::hip] ") + what + ": " + hipGetErrorString(rc);
  return false;
}

// The status strings of the engine an engine stands in for: the SIMD
// engine's (simd...:381-422, :703-780) or the scalar engine's
// (sisd_alignment_engine.cpp:62-77, :255-283).
struct EngineMsgs {
  const char* too_large;
  const char* overflow;
  const char* no_memory;
  const char* pre_too_large;
  const char* pre_no_memory;
};
inline constexpr EngineMsgs kSimdMsgs{
    "[spoa::SimdAlignmentEngine::Align] error: too large sequence!",
    "[spoa::SimdAlignmentEngine::Align] error: possible overflow!",
    "[spoa::SimdAlignmentEngine::Align] error: insufficient memory!",
    "[spoa::SimdAlignmentEngine::Prealloc] error: too large sequence!",
    "[spoa::SimdAlignmentEngine::Prealloc] error: insufficient memory!"};
inline constexpr EngineMsgs kSisdMsgs{
    "[spoa::SisdAlignmentEngine::Align] error: too large sequence!",
    "[spoa::SisdAlignmentEngine::Align] error: possible overflow!",
    "[spoa::SisdAlignmentEngine::Align] error: insufficient memory!",
    "[spoa::SisdAlignmentEngine::Prealloc] error: too large sequence!",
    "[spoa::SisdAlignmentEngine::Prealloc] error: insufficient memory!"};

constexpr uint64_t kAlign = 256;
inline uint64_t RoundUp(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// reference src/alignment_engine.cpp:101-110
inline int64_t WorstCase(const spoa_hip_params& p, int64_t i, int64_t j) {
  auto gap = [&](int64_t len) -> int64_t {
    return len == 0 ? 0 : std::min<int64_t>(p.g + (len - 1) * p.e, p.q + (len - 1) * p.c);
  };
  const int64_t d = i > j ? i - j : j - i;
  return std::min<int64_t>(-1 * (p.m * std::min(i, j) + gap(d)), gap(i) + gap(j));
}

// Score width for one call (reference simd...:719-774): 0 means "possible
// overflow!".
inline int WidthBits(const spoa_hip_params& p, uint32_t len, uint32_t V) {
  const int64_t w = WorstCase(p, static_cast<int64_t>(len) + 8, V);
  if (w < std::numeric_limits<int32_t>::min() + 1024) return 0;
  if (w < std::numeric_limits<int16_t>::min() + 1024) return 32;
  return 16;
}

inline int Lanes(int width_bits) { return width_bits == 16 ? 16 : 8; }

// How one job runs: `width` is the reference's score width (what the caller
// is told), `bits` the width of the stored rows, `lanes` the reference lane
// count whose row width and convex segments the DP reproduces.  Local
// alignments that the reference runs at int32 keep int16 rows when every
// score fits them: H stays in [0, m * n] and every gap score is at least
// its penalty above a non-negative H, so only m * (n + 8) needs headroom
// (kept at 30000 below int16's maximum).
struct Shape {
  int width = 0, bits = 0, lanes = 0;
  int cfgw = 0;  // the width whose launch configuration the job takes
};

// With `sisd` the job reproduces the reference's scalar engine
// (sisd_alignment_engine.cpp): its overflow bound is on (n, V) and it
// scores in int32 (the width reported); its rows have no padding and its
// convex E / Q open from the final H of every column, i.e. one-column
// segments (lanes = 1).  Its scores are those of the SIMD engine's recurrences,
// so they fit int16 rows wherever the SIMD engine's int16 bound holds (or
// the local-alignment rule above), and such jobs take the int16 launch
// configuration.
inline Shape JobShape(const spoa_hip_params& p, uint32_t len, uint32_t V, bool sisd) {
  Shape sh;
  const bool sw16 = p.type == SPOA_HIP_SW &&
                    static_cast<int64_t>(std::max<int32_t>(p.m, 0)) * (static_cast<int64_t>(len) + 8) <= 30000;
  if (sisd) {
    if (WorstCase(p, len, V) < std::numeric_limits<int32_t>::min() + 1024) return sh;
    sh.width = 32;
    sh.lanes = 1;
    sh.bits = WidthBits(p, len, V) == 16 || sw16 ? 16 : 32;
    sh.cfgw = sh.bits;
    return sh;
  }
  sh.width = WidthBits(p, len, V);
  if (sh.width == 0) return sh;
  sh.lanes = Lanes(sh.width);
  sh.bits = sh.width;
  sh.cfgw = sh.width;
  if (sh.width == 32 && sw16) sh.bits = 16;
  return sh;
}
inline int Matrices(int subtype) { return subtype == 0 ? 1 : (subtype == 1 ? 3 : 5); }
inline int ReadMatrices(int subtype) { return subtype == 0 ? 1 : (subtype == 1 ? 2 : 3); }

// Launch configuration: columns per lane, waves per workgroup, and whether
// a job whose row spans several tiles of waves x 64 x k columns runs each
// tile on its own workgroup (split) or all of them in turn on one.
struct Cfg {
  int k = 8;
  int waves = 1;
  int split = 0;
  bool operator<(const Cfg& o) const { return std::tie(k, waves, split) < std::tie(o.k, o.waves, o.split); }
};

// Row tiles of a job with row stride `stride` under cfg that run on
// workgroups of their own: 1 unless the job is split.
inline uint32_t JobTiles(uint32_t stride, const Cfg& cfg) {
  if (!cfg.split) return 1;
  const uint32_t tile = 64u * static_cast<uint32_t>(cfg.k) * static_cast<uint32_t>(cfg.waves);
  return (stride + tile - 1) / tile;
}

// A split job's 64-bit hand-off words (spoa_hip_dev::DevJob): three per row
// and tile boundary, two per tile.
inline uint64_t XferWords(uint32_t V, uint32_t tiles) {
  if (tiles <= 1) return 0;
  const uint64_t w = 3ull * (tiles - 1) * (static_cast<uint64_t>(V) + 1) + 2ull * tiles;
  return (w + 1) / 2 * 2;
}

// Jobs are split into at most this many tiles (longer rows run their tiles
// in turn).
constexpr uint32_t kMaxSplitTiles = 4096;

// Device buffer, grow-only.
struct DevBuf {
  void* p = nullptr;
  uint64_t size = 0;
  bool Reserve(uint64_t need, const char* what) {
    if (need <= size) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    size = 0;
    const uint64_t want = std::max<uint64_t>(need, std::min<uint64_t>(2 * need, std::max<uint64_t>(size * 2, need + need / 4)));
    if (hipMalloc(&p, want) == hipSuccess) {
      size = want;
      return true;
    }
    (void)hipGetLastError();
    if (hipMalloc(&p, need) == hipSuccess) {
      size = need;
      return true;
    }
    (void)hipGetLastError();
    p = nullptr;
    g_last_error = what;
    return false;
  }
  void Release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    size = 0;
  }
};

// Pinned host buffer, grow-only.
struct PinBuf {
  void* p = nullptr;
  uint64_t size = 0;
  bool Reserve(uint64_t need) {
    if (need <= size) return true;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    size = 0;
    const uint64_t want = std::max<uint64_t>(need, 2 * need);
    if (!HipOk(hipHostMalloc(&p, want, hipHostMallocDefault), "hipHostMalloc")) return false;
    size = want;
    return true;
  }
  void Release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    size = 0;
  }
  template <typename T>
  T* At(uint64_t off) {
    return reinterpret_cast<T*>(static_cast<uint8_t*>(p) + off);
  }
};

using KernelFn = void (*)(KArgs);

// Multi-wave workgroups have at most 8 waves (two a SIMD), which leaves
// each lane 256 registers.
constexpr int kMaxGroupWaves = 8;

// max-plus 2x2 product
inline void MpMul(const int32_t* a, const int32_t* b, int32_t* out) {
  int32_t r[4];
  r[0] = std::max(a[0] + b[0], a[1] + b[2]);
  r[1] = std::max(a[0] + b[1], a[1] + b[3]);
  r[2] = std::max(a[2] + b[0], a[3] + b[2]);
  r[3] = std::max(a[2] + b[1], a[3] + b[3]);
  std::memcpy(out, r, sizeof(r));
}

inline KParams MakeParams(const spoa_hip_params& p, int width_bits, int lanes, int k, bool sisd) {
  KParams kp{};
  kp.sisd = sisd ? 1 : 0;
  kp.type = p.type;
  kp.m = p.m;
  kp.mis = p.n;
  kp.g = p.g;
  kp.e = p.e;
  kp.q = p.q;
  kp.c = p.c;
  // reference simd...:526-528
  kp.pad = -1 * std::max(std::max(std::abs(static_cast<int>(p.m)), std::abs(static_cast<int>(p.n))),
                         std::max(std::abs(static_cast<int>(p.g)), std::abs(static_cast<int>(p.q))));
  kp.ninf = width_bits == 16 ? std::numeric_limits<int16_t>::min() + 1024
                             : std::numeric_limits<int32_t>::min() + 1024;
  kp.decay = p.subtype == 0 ? k * p.g : k * p.e;
  if (p.subtype == 2) {
    // one L-column segment: (E, Q) entering the segment -> entering the next
    const int L = lanes;
    const int32_t seg[4] = {(L - 1) * p.e + std::max<int32_t>(p.g, p.e), (L - 1) * p.c + p.g,
                            (L - 1) * p.e + p.q, (L - 1) * p.c + std::max<int32_t>(p.q, p.c)};
    std::memcpy(kp.mseg, seg, sizeof(seg));
    int32_t lane[4];
    std::memcpy(lane, seg, sizeof(seg));
    for (int s = 1; s < k / L; ++s) MpMul(seg, lane, lane);
    std::memcpy(kp.mpow[0], lane, sizeof(lane));
    for (int t = 1; t < 7; ++t) MpMul(kp.mpow[t - 1], kp.mpow[t - 1], kp.mpow[t]);
  }
  return kp;
}

// The alignment kernel of a launch group (defined in poa_hip.hip):
// subtype, stored row width, reference lane count, columns per lane, waves,
// whether the launch may hold split jobs and whether it reproduces the
// scalar engine.
KernelFn PickKernel(int subtype, int bits, int lanes, int k, int waves, bool split, bool sisd);

}  // namespace spoa_hip_host

using spoa_hip_host::Cfg;
using spoa_hip_host::DevBuf;
using spoa_hip_host::KernelFn;
using spoa_hip_host::PinBuf;

struct Plan {
  uint32_t job;
  int width;  // the reference's score width
  int bits;   // stored rows
  int lanes;  // reference lane count (row width, convex segments)
  Cfg cfg;
  uint32_t W, stride;
  uint64_t mat_bytes;
  uint64_t graph_words;
  uint64_t seq_bytes;
  uint64_t work;  // V * n: the queue order inside a launch group
  uint32_t tiles = 1;  // workgroups of a split job (queue items), else 1
};

struct spoa_hip_engine {
  int device = 0;
  // host threads packing and unpacking large passes (made on first use)
  std::unique_ptr<spoa_amd::ThreadPool> pool;
  hipStream_t stream = nullptr;   // the engine's own stream
  hipStream_t copy = nullptr;     // copies: the own stream, or a shared one
  // kernels: the engine's own stream, or one shared with other engines so
  // that their launches run one after another and never share the device
  hipStream_t compute = nullptr;
  hipStream_t compute_own = nullptr;  // created when other engines share this one's kernels
  // launch groups after a pass's first run on this stream, beside the first
  // (for example the few int32-width jobs of a pass beside its int16 jobs):
  // the engine's own stream when its kernels run on a shared compute stream
  // (its copies are idle while its pass runs), else a stream made on first
  // use (streams map onto the device's few hardware queues in creation
  // order, and engines that never need one must not shift that mapping for
  // the streams made after them)
  hipStream_t side = nullptr;
  hipStream_t side_own = nullptr;
  hipEvent_t ev_side_in = nullptr, ev_side = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;  // copy timing
  hipEvent_t ev_in = nullptr, ev_kdone = nullptr;  // copy -> kernels -> copy ordering
  std::vector<hipEvent_t> ev_ks, ev_ke;            // start / end of every launch of a pass
  size_t pass_launches = 0;
  // graph mode: every pass's launch sequence is captured into a HIP graph
  // and launched as one unit on the compute stream
  bool graph_mode = false;
  hipGraphExec_t graph_exec = nullptr;  // kept from pass to pass, updated in place
  bool graph_pass = false;  // the pending pass ran as a graph
  std::vector<int> pass_width;  // score width of every launch of the pass
  spoa_hip_params prm{};
  // the reference's scalar engine's results and messages instead of its
  // SIMD engine's (spoa_hip_engine_create_sisd)
  bool sisd = false;
  const spoa_hip_host::EngineMsgs* msg = &spoa_hip_host::kSimdMsgs;
  // row-phase cycle counters of the alignment kernels (builds with
  // SPOA_ROW_PHASES; null otherwise)
  unsigned long long* phases = nullptr;
  int num_cus = 256;
  uint64_t budget = 0;
  DevBuf d_in, d_out, d_counter;
  // split jobs: their hand-off words (zeroed when allocated; each launch
  // tags its words with a new epoch), and the longest wait of one tile for
  // another (100 MHz clock ticks) before the job reports a time-out
  DevBuf d_xfer;
  uint32_t epoch = 0;
  uint32_t wait_ticks = 1000000000u;
  // Buffers of the last device-resident window set the engine ran
  // (spoa_hip_windows_destroy hands them back, the next
  // spoa_hip_windows_create takes them over), so that repeated runs of
  // window sets allocate device and pinned memory only to grow it.
  struct WinBuffers {
    DevBuf d[12];
    PinBuf h[5];
    std::vector<hipEvent_t> ev_s, ev_e;
    hipEvent_t ev_m0 = nullptr, ev_m1 = nullptr, ev_done = nullptr;
    bool held = false;
  } win_bufs;
  // DP matrices: grow-only segments kept for the engine's lifetime, so a
  // long-lived engine allocates once (allocation time grows with size)
  std::vector<DevBuf> arena;
  uint64_t arena_bytes = 0;
  // a launched pass not yet finished (the asynchronous batch form)
  struct Pending {
    bool active = false;
    spoa_hip_job* jobs = nullptr;
    std::vector<Plan> plans;
    size_t lo = 0, hi = 0;
    uint64_t off_res = 0, off_pairs = 0;
    std::vector<uint64_t> pair_at;
    int first = 0;  // status of the jobs rejected before any launch
  } pending;
  PinBuf h_in, h_out;
  spoa_hip_stats stats{};
  std::map<std::tuple<KernelFn, int, uint32_t>, int> occupancy;  // (kernel, threads, LDS) -> blocks a CU
  // overrides: (width, batch) -> max_width -> cfg
  std::map<std::pair<int, int>, std::map<uint32_t, Cfg>> overrides;
};

namespace spoa_hip_host {

// Frees the window-set buffers and events an engine keeps (WinBuffers).
inline void ReleaseWinBuffers(spoa_hip_engine* E) {
  auto& c = E->win_bufs;
  for (auto& b : c.d) b.Release();
  for (auto& b : c.h) b.Release();
  for (auto ev : c.ev_s) (void)hipEventDestroy(ev);
  for (auto ev : c.ev_e) (void)hipEventDestroy(ev);
  c.ev_s.clear();
  c.ev_e.clear();
  for (hipEvent_t* ev : {&c.ev_m0, &c.ev_m1, &c.ev_done}) {
    if (*ev) (void)hipEventDestroy(*ev);
    *ev = nullptr;
  }
  c.held = false;
}


// Batched launches: 16 columns a lane for int16 rows (at most 1,024
// columns a wave), 8 for int32 rows (512 a wave), as many waves as the row
// needs up to 8 (wider rows take several tiles).  Measured on batches of
// 1.5 kb local convex and 500 bp overlap affine windows, this beats both
// longer lanes (K = 32, one wave) and more waves of shorter lanes.  The per-call launch takes the widest workgroup
// (up to 8 waves) of the shortest lanes.
inline Cfg DefaultCfg(const spoa_hip_engine* E, uint32_t W, int width_bits, bool batch) {
  Cfg c;
  const bool convex16 = E->prm.subtype == 2 && width_bits == 16;
  const int kmin = convex16 ? 16 : 8;
  c.k = batch ? (width_bits == 16 ? 16 : 8) : kmin;
  const uint32_t per_wave = 64u * static_cast<uint32_t>(c.k);
  c.waves = static_cast<int>(
      std::min<uint32_t>(kMaxGroupWaves, std::max<uint32_t>(1u, (W + per_wave - 1) / per_wave)));
  // Rows wider than one workgroup of 8 waves: every one-wave tile of the
  // row on its own workgroup (no barriers inside a tile, and the tiles of
  // one job run side by side on as many SIMDs).
  if (W > per_wave * kMaxGroupWaves && (W + per_wave - 1) / per_wave <= kMaxSplitTiles) {
    c.waves = 1;
    c.split = 1;
  }
  return c;
}

inline Cfg ChooseCfg(const spoa_hip_engine* E, uint32_t len, int width_bits, bool batch) {
  const int L = Lanes(width_bits);
  const uint32_t W = (len + L - 1) / L * L;
  auto it = E->overrides.find({width_bits, batch ? 1 : 0});
  if (it != E->overrides.end()) {
    auto jt = it->second.lower_bound(W);
    if (jt != it->second.end()) return jt->second;
  }
  return DefaultCfg(E, W, width_bits, batch);
}

// Device bytes of one job: its DP matrices, plus -- when a row spans
// several wave segments (waves x tiles) -- the H value left of every
// segment for every row (int32), which the first lane of a segment keeps
// for the diagonal moves of later rows.
inline uint64_t JobBytes(const spoa_hip_params& p, uint32_t V, uint32_t stride, int width, const Cfg& cfg) {
  uint64_t b = static_cast<uint64_t>(V + 1) * Matrices(p.subtype) * stride * (width / 8);
  const uint64_t tile = 64ull * static_cast<uint64_t>(cfg.k) * static_cast<uint64_t>(cfg.waves);
  const uint64_t nseg = (stride + tile - 1) / tile * static_cast<uint64_t>(cfg.waves);
  if (nseg > 1) b += nseg * (V + 1) * 4ull;
  return RoundUp(b, kAlign);
}

inline KernelFn GroupKernel(spoa_hip_engine* E, const Plan& pl) {
  return PickKernel(E->prm.subtype, pl.bits, pl.lanes, pl.cfg.k, pl.cfg.waves, pl.cfg.split != 0, E->sisd);
}

// Launch group of a job: jobs of one group share a kernel and its
// parameters.  Linear and affine rows depend on the lane count only through
// the row width each job carries, so their int16 rows share one group.
inline int GroupLanes(int subtype, const Plan& pl) { return subtype == 2 || pl.bits == 32 ? pl.lanes : 16; }


inline bool SameGroup(int subtype, const Plan& a, const Plan& b) {
  return a.bits == b.bits && GroupLanes(subtype, a) == GroupLanes(subtype, b) && a.cfg.k == b.cfg.k &&
         a.cfg.waves == b.cfg.waves && a.cfg.split == b.cfg.split;
}

// One persistent launch of a pass: the jobs [a, b) of the pass that share a
// kernel, with its launch shape and parameters.
struct LaunchGroup {
  size_t a, b;
  uint32_t items;       // queue items (the tiles of its split jobs count each)
  uint64_t item_at;     // its first item in the pass's item table
  KernelFn fn;
  int threads;
  uint32_t lds;   // query-profile bytes (4 bases x K/2 words a lane), or more
  uint32_t grid;
  uint32_t prio;  // leading jobs (largest V * n) at raised wave priority
  KParams prm;
};

// Zeroes n launch-group job counters on a stream (a kernel, so that the
// sequence stays capturable); defined in poa_hip.hip.
bool LaunchResetCounters(uint32_t* counters, uint32_t n, hipStream_t stream);

// Places matrices of the given sizes in the engine's arena (first fit,
// growing it within the budget by segments of at least min_segment bytes);
// defined in poa_hip.hip.
constexpr uint64_t kBatchSegment = 4ull << 30;      // batched passes
constexpr uint64_t kPerCallSegment = 64ull << 20;   // one alignment per call
int PlaceMatrices(spoa_hip_engine* E, const std::vector<uint64_t>& need, std::vector<uint64_t>* addr,
                  uint64_t min_segment);

// Split jobs (defined in poa_hip.hip): the engine's hand-off words grown to
// at least `words` (zeroed when allocated; no launch of the engine may be
// running), and the next launch's tag.
bool ReserveXfer(spoa_hip_engine* E, uint64_t words);
uint32_t NextEpoch(spoa_hip_engine* E);
inline constexpr const char* kSplitTimeoutMsg =
    "[spoa::hip] the row tiles of a split alignment timed out waiting for each other";

// Launch groups of plans[lo, hi) (sorted by group, largest V * n first in
// one); defined in poa_hip.hip.
int BuildGroups(spoa_hip_engine* E, const std::vector<Plan>& plans, size_t lo, size_t hi,
                std::vector<LaunchGroup>* out);

}  // namespace spoa_hip_host
