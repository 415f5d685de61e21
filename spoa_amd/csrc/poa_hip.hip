// Host half of the C ABI in include/spoa_hip.h: device buffers, job
// packing, launch configuration and the per-call / batched launches of the
// alignment kernels in poa_kernels.hpp.
//
// Per call (AlignmentEngine::Align) one job is packed into pinned memory
// (rank-ordered CSR, first columns, read), copied with one transfer,
// aligned by one workgroup of up to 8 waves, and its traceback pairs come
// back with one transfer.  A batch is grouped by launch configuration,
// ordered by V * n (largest first, the queue order of the persistent
// grid) and split into launches whose DP matrices fit the memory budget.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "engine.hpp"
#include "pass_split.hpp"
#include "pick_kernel.hpp"
#include "spoa_hip.h"
#include "thread_pool.hpp"

using namespace spoa_hip_host;

namespace spoa_hip_host {

// bits: stored rows; lanes: the reference lane count the launch reproduces.
// Only convex rows depend on the lane count beyond the row width each job
// carries, so linear and affine rows always run the kernel of their width's
// own lane count; convex int16 rows with 8 lanes hold the local alignments
// the reference scores at int32, and one-column segments are the scalar
// engine's convex rows.  split: the launch may hold split jobs; sisd: it
// reproduces the scalar engine.  The families are built in poa_family.hip.
KernelFn PickKernel(int subtype, int bits, int lanes, int k, int waves, bool split, bool sisd) {
  switch (subtype) {
    case 0:
      return bits == 16 ? PickFamily<kFamLinear16>(bits, lanes, k, waves, split, sisd)
                        : PickFamily<kFamLinear32>(bits, lanes, k, waves, split, sisd);
    case 1:
      return bits == 16 ? PickFamily<kFamAffine16>(bits, lanes, k, waves, split, sisd)
                        : PickFamily<kFamAffine32>(bits, lanes, k, waves, split, sisd);
    default:
      if (lanes == 1) return PickFamily<kFamConvexGotoh>(bits, lanes, k, waves, split, sisd);
      if (bits == 32) return PickFamily<kFamConvex32>(bits, lanes, k, waves, split, sisd);
      return PickFamily<kFamConvex16>(bits, lanes, k, waves, split, sisd);
  }
}


}  // namespace spoa_hip_host

namespace spoa_hip_dev {

// Zeroes the job counters of a pass (a kernel rather than a memset so the
// launch sequence can be captured in a graph).
__global__ void ResetCounters(uint32_t* p, uint32_t n) {
  if (blockIdx.x == 0 && threadIdx.x < n) p[threadIdx.x] = 0u;
}

}  // namespace spoa_hip_dev

namespace spoa_hip_host {

bool LaunchResetCounters(uint32_t* counters, uint32_t n, hipStream_t stream) {
  spoa_hip_dev::ResetCounters<<<1, 256, 0, stream>>>(counters, n);
  return HipOk(hipGetLastError(), "counter reset");
}

// Places matrices of the given sizes in the engine's arena segments (first
// fit, in order), growing the arena within the memory budget (at least
// doubling it, so that batches whose jobs grow round by round allocate
// O(log) times); when the segments are too fragmented, starts over with one
// segment holding everything.  Device addresses into addr.
int PlaceMatrices(spoa_hip_engine* E, const std::vector<uint64_t>& need, std::vector<uint64_t>* addr_out,
                  uint64_t min_segment) {
  const size_t nj = need.size();
  std::vector<uint64_t>& addr = *addr_out;
  addr.assign(nj, 0);
  std::vector<uint64_t> used(E->arena.size(), 0);
  size_t x = 0;
  bool retried = false;
  while (x < nj) {
    bool placed = false;
    for (size_t sgi = 0; sgi < E->arena.size(); ++sgi) {
      if (used[sgi] + need[x] <= E->arena[sgi].size) {
        addr[x] = reinterpret_cast<uint64_t>(E->arena[sgi].p) + used[sgi];
        used[sgi] += need[x];
        placed = true;
        break;
      }
    }
    if (placed) {
      ++x;
      continue;
    }
    uint64_t rest = 0;
    for (size_t y = x; y < nj; ++y) rest += need[y];
    const uint64_t room = E->budget > E->arena_bytes ? E->budget - E->arena_bytes : 0;
    uint64_t want = std::min(room, std::max<uint64_t>(std::max<uint64_t>(rest, min_segment), E->arena_bytes));
    if (want < need[x]) {
      // fragmented: start over with one segment of twice what the whole
      // set needs (within the budget), so that sets which grow round by
      // round do not free and allocate the arena every round
      if (retried) return Fail(SPOA_HIP_NO_MEMORY, E->msg->no_memory);
      retried = true;
      for (auto& sg : E->arena) sg.Release();
      E->arena.clear();
      E->arena_bytes = 0;
      used.clear();
      uint64_t all = 0;
      for (uint64_t b : need) all += b;
      DevBuf sg;
      const uint64_t one = std::min<uint64_t>(E->budget, std::max<uint64_t>(2 * all, min_segment));
      if (one >= all && hipMalloc(&sg.p, one) == hipSuccess) {
        sg.size = one;
        E->arena.push_back(sg);
        E->arena_bytes += one;
        used.push_back(0);
      } else {
        (void)hipGetLastError();
      }
      x = 0;
      continue;
    }
    DevBuf sg;
    if (hipMalloc(&sg.p, want) != hipSuccess) {
      (void)hipGetLastError();
      want = rest;
      if (want < need[x] || hipMalloc(&sg.p, want) != hipSuccess) {
        (void)hipGetLastError();
        return Fail(SPOA_HIP_NO_MEMORY, E->msg->no_memory);
      }
    }
    sg.size = want;
    E->arena.push_back(sg);
    E->arena_bytes += want;
    used.push_back(0);
  }
  return SPOA_HIP_OK;
}

// The hand-off words of a pass's split jobs: grown (and zeroed, so that no
// word carries a tag before some launch writes it) while no launch of the
// engine uses them.
bool ReserveXfer(spoa_hip_engine* E, uint64_t words) {
  const uint64_t bytes = words * 8;
  if (E->d_xfer.size >= bytes) return true;
  if (!E->d_xfer.Reserve(bytes, E->msg->no_memory)) return false;
  return HipOk(hipMemsetAsync(E->d_xfer.p, 0, E->d_xfer.size, E->copy), "hipMemsetAsync") &&
         HipOk(hipStreamSynchronize(E->copy), "hipStreamSynchronize");
}

// The tag of a launch's hand-off words: never 0 (the zeroed words).
uint32_t NextEpoch(spoa_hip_engine* E) {
  if (++E->epoch == 0) E->epoch = 1;
  return E->epoch;
}

// Launch groups of a pass: plans[lo, hi) sorted by launch group, the largest
// V * n first inside a group.  Per group: its kernel, workgroup size, LDS,
// grid (workgroups that fit the device at once, at most one a job) and the
// number of leading jobs that run at raised wave priority.
int BuildGroups(spoa_hip_engine* E, const std::vector<Plan>& plans, size_t lo, size_t hi,
                std::vector<LaunchGroup>* out) {
  const spoa_hip_params& P = E->prm;
  std::vector<LaunchGroup>& groups = *out;
  groups.clear();
  uint64_t item_at = 0;
  for (size_t x = lo; x < hi;) {
    size_t y = x;
    while (y < hi && SameGroup(E->prm.subtype, plans[y], plans[x])) ++y;
    LaunchGroup gp;
    gp.a = x - lo;
    gp.b = y - lo;
    uint64_t items = 0;
    for (size_t z = x; z < y; ++z) items += plans[z].tiles;
    if (items > 0xffffffffull) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] too many row tiles in one launch");
    gp.items = static_cast<uint32_t>(items);
    gp.item_at = item_at;
    item_at += items;
    gp.fn = GroupKernel(E, plans[x]);
    gp.threads = 64 * plans[x].cfg.waves;
    gp.lds = 8u * static_cast<uint32_t>(plans[x].cfg.k) * static_cast<uint32_t>(gp.threads);
    if (!gp.fn) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] unsupported launch configuration");
    auto blocks_per_cu = [&](uint32_t lds, int* n_out) -> bool {
      auto key = std::make_tuple(gp.fn, gp.threads, lds);
      auto oc = E->occupancy.find(key);
      if (oc != E->occupancy.end()) {
        *n_out = oc->second;
        return true;
      }
      if (!HipOk(hipOccupancyMaxActiveBlocksPerMultiprocessor(n_out, gp.fn, gp.threads, lds), "occupancy")) {
        return false;
      }
      E->occupancy[key] = *n_out;
      return true;
    };
    int per_cu = 0;
    if (!blocks_per_cu(gp.lds, &per_cu)) return SPOA_HIP_DEVICE_ERROR;
    const uint32_t n = gp.items;
    // Workgroups take the LDS of their query profile (the traceback's tile
    // lives in it), enlarged to 24 or 32 KiB only where that leaves the
    // kernel's own workgroups per CU (set by its registers) unchanged:
    // launches of the two halves and of several launch groups share the
    // device, and workgroups enlarged to fill it alone kept the other
    // launches' workgroups waiting (11% more windows/s on 2,048 local
    // convex 1.5 kb windows without that enlargement; about 1% more with
    // this one).
    for (uint32_t cand : {32768u, 24576u}) {
      if (cand <= gp.lds) break;
      int pc = 0;
      if (!blocks_per_cu(cand, &pc)) return SPOA_HIP_DEVICE_ERROR;
      if (pc >= per_cu) {
        gp.lds = cand;
        break;
      }
    }
    gp.grid = std::max<uint32_t>(
        1u, std::min<uint32_t>(n, static_cast<uint32_t>(std::max(1, per_cu)) * static_cast<uint32_t>(E->num_cus)));
    gp.prm = MakeParams(P, plans[x].bits, plans[x].lanes, plans[x].cfg.k, E->sisd);
    // the group's jobs of at least half its largest job's V * n (they lead
    // the queue) get the raised wave priority (1.7% less kernel time on
    // 2,048 local convex 1.5 kb windows; raising the side launches' jobs
    // above them was slower)
    const uint64_t top = plans[x].work;
    gp.prio = 0;
    while (x + gp.prio < y && 2 * plans[x + gp.prio].work >= top) ++gp.prio;
    groups.push_back(gp);
    x = y;
  }
  return SPOA_HIP_OK;
}

}  // namespace spoa_hip_host

namespace {

// First columns (reference simd...:568-577, 590-599, 607-694): fcF / fcO
// are the vertical-gap columns, fcH the primary column per alignment type.
// Local and overlap alignments start every row at 0 and never walk the
// first column back (the traceback stops at column 0), so there only the
// zero H column matters and the vertical-gap columns stay zero.
void FirstColumns(const spoa_hip_params& p, const spoa_hip_graph& gr, int32_t* fcH, int32_t* fcF, int32_t* fcO) {
  const uint32_t V = gr.num_nodes;
  if (p.type != SPOA_HIP_NW) {
    std::memset(fcH, 0, (V + 1) * sizeof(int32_t));
    std::memset(fcF, 0, (V + 1) * sizeof(int32_t));
    std::memset(fcO, 0, (V + 1) * sizeof(int32_t));
    return;
  }
  const int32_t kNinf = std::numeric_limits<int32_t>::min() + 1024;
  fcF[0] = 0;
  fcO[0] = 0;
  for (uint32_t i = 1; i <= V; ++i) {
    const uint32_t b = gr.pred_off[i - 1], e = gr.pred_off[i];
    int32_t pf = b == e ? p.g - p.e : kNinf;
    int32_t po = b == e ? p.q - p.c : kNinf;
    for (uint32_t k = b; k < e; ++k) {
      const uint32_t row = gr.pred_rank[k] + 1;
      pf = std::max(pf, fcF[row]);
      po = std::max(po, fcO[row]);
    }
    fcF[i] = pf + p.e;
    fcO[i] = po + p.c;
  }
  if (p.type == SPOA_HIP_NW) {
    if (p.subtype == 2) {
      for (uint32_t i = 0; i <= V; ++i) fcH[i] = std::max(fcF[i], fcO[i]);
    } else if (p.subtype == 1) {
      for (uint32_t i = 0; i <= V; ++i) fcH[i] = fcF[i];
    } else {
      fcH[0] = 0;
      for (uint32_t i = 1; i <= V; ++i) {
        const uint32_t b = gr.pred_off[i - 1], e = gr.pred_off[i];
        int32_t ph = b == e ? 0 : kNinf;
        for (uint32_t k = b; k < e; ++k) ph = std::max(ph, fcH[gr.pred_rank[k] + 1]);
        fcH[i] = ph + p.g;
      }
    }
  } else {
    for (uint32_t i = 0; i <= V; ++i) fcH[i] = 0;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// The entry point that allocates nothing (spoa_hip_align_noalloc): every
// device byte it uses beyond the caller's inputs and outputs -- the job
// record, the packed graph block, the padded read, the DP matrices, the
// queue counter and the kernel's result record -- lives in this one
// file-scope array.  Each GPU holds its own copy from the first use of the
// library on it until the process ends; every call on one GPU shares it, so
// calls that overlap on one GPU race.  Host code never passes the array's
// name to a kernel (on the host it is only a stand-in): the device address
// of each GPU's copy comes from hipGetSymbolAddress at engine creation.
constexpr uint64_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch[kMaxDevices] = {};

// The scratch array's device address on `device`, which must be current.
int ScratchInit(int device) {
  if (device < 0 || device >= kMaxDevices) return Fail(SPOA_HIP_DEVICE_ERROR, "[spoa::hip] device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return SPOA_HIP_OK;
  void* ptr = nullptr;
  if (!HipOk(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) {
    return SPOA_HIP_DEVICE_ERROR;
  }
  g_scratch[device] = static_cast<uint8_t*>(ptr);
  return SPOA_HIP_OK;
}

}  // namespace

namespace spoa_hip_dev {

// Arguments of the packing kernel of the no-allocation entry point.
struct NoAllocPackArgs {
  // the caller's rank view and read (device memory)
  const uint8_t* row_char;
  const int32_t* node_id;
  const uint32_t* pred_off;
  const uint32_t* pred_rank;
  const uint8_t* is_sink;
  const char* seq;
  uint32_t V, n, W, stride, cap;
  int32_t type, subtype, g, e, q, c;
  // inside the scratch array
  uint32_t* counter;
  DevJob* job;
  int32_t* res;       // the kernel's result record: n_pairs, score, has_score, status
  uint8_t* seq_out;   // stride + 16 bytes
  uint64_t mat_addr;  // DP matrices
  uint32_t* graph;    // graph block (layout of MakeView)
  uint64_t graph_cap; // 32-bit words available for the graph block
};

// Builds the job record, the graph block (row records, node ids, predecessor
// offsets, first columns, predecessor rows) and the zero-padded read in the
// scratch array, and arms the queue counter: 0 when the job fits, 1 (the
// alignment kernel then takes no job) with status "insufficient memory"
// when its predecessor lists do not fit.  First columns as FirstColumns on
// the host (reference simd...:568-577, 590-599, 607-694): one thread walks
// the ranks in order.  One workgroup.
__global__ void NoAllocPack(NoAllocPackArgs a) {
  const uint32_t V = a.V;
  const uint32_t base = a.pred_off[0];
  const uint64_t E = static_cast<uint64_t>(a.pred_off[V]) - base;
  const bool fits = 9ull * V + 4 + E <= a.graph_cap;
  if (!fits) {
    if (threadIdx.x == 0) {
      *a.counter = 1u;
      a.res[0] = 0;
      a.res[1] = 0;
      a.res[2] = 0;
      a.res[3] = SPOA_HIP_NO_MEMORY;
    }
    return;
  }
  uint32_t* rows = a.graph;
  int32_t* node_id = reinterpret_cast<int32_t*>(a.graph + 4ull * V);
  uint32_t* poff = a.graph + 5ull * V;
  int32_t* fcH = reinterpret_cast<int32_t*>(a.graph + 6ull * V + 1);
  int32_t* fcF = fcH + (V + 1);
  int32_t* fcO = fcF + (V + 1);
  uint32_t* preds = a.graph + 9ull * V + 4;
  if (threadIdx.x == 0) {
    const int32_t kNinf = -2147483647 - 1 + 1024;
    if (a.type != SPOA_HIP_NW) {
      for (uint32_t i = 0; i <= V; ++i) fcH[i] = fcF[i] = fcO[i] = 0;
    } else {
      fcF[0] = 0;
      fcO[0] = 0;
      for (uint32_t i = 1; i <= V; ++i) {
        const uint32_t b = a.pred_off[i - 1], e = a.pred_off[i];
        int32_t pf = b == e ? a.g - a.e : kNinf;
        int32_t po = b == e ? a.q - a.c : kNinf;
        for (uint32_t k = b; k < e; ++k) {
          const uint32_t row = a.pred_rank[k] + 1;
          pf = max(pf, fcF[row]);
          po = max(po, fcO[row]);
        }
        fcF[i] = pf + a.e;
        fcO[i] = po + a.c;
      }
      if (a.subtype == 2) {
        for (uint32_t i = 0; i <= V; ++i) fcH[i] = max(fcF[i], fcO[i]);
      } else if (a.subtype == 1) {
        for (uint32_t i = 0; i <= V; ++i) fcH[i] = fcF[i];
      } else {
        fcH[0] = 0;
        for (uint32_t i = 1; i <= V; ++i) {
          const uint32_t b = a.pred_off[i - 1], e = a.pred_off[i];
          int32_t ph = b == e ? 0 : kNinf;
          for (uint32_t k = b; k < e; ++k) ph = max(ph, fcH[a.pred_rank[k] + 1]);
          fcH[i] = ph + a.g;
        }
      }
    }
    DevJob j{};
    j.mat_addr = a.mat_addr;
    j.graph_off = 0;
    j.seq_off = 0;
    j.V = V;
    j.n = a.n;
    j.W = a.W;
    j.stride = a.stride;
    j.pair_off = 0;
    j.cap = a.cap;
    j.nslots = 0;  // the query profile is formed per row
    j.slot_chars = 0;
    *a.job = j;
    *a.counter = 0u;
    a.res[0] = 0;
    a.res[1] = 0;
    a.res[2] = 0;
    a.res[3] = 0;
  }
  __syncthreads();
  for (uint32_t r = threadIdx.x; r < V; r += blockDim.x) {
    const uint32_t pb = a.pred_off[r] - base, pe = a.pred_off[r + 1] - base;
    const uint32_t npred = pe > pb ? pe - pb : 1u;
    rows[4ull * r + 0] = static_cast<uint32_t>(a.row_char[r]) | (a.is_sink[r] ? 0x100u : 0u) | (7u << 9) |
                         (npred << 12);
    rows[4ull * r + 1] = pe > pb ? a.pred_rank[a.pred_off[r]] + 1 : 0u;
    rows[4ull * r + 2] = pb;
    rows[4ull * r + 3] = static_cast<uint32_t>(fcH[r + 1]);
    node_id[r] = a.node_id[r];
    poff[r] = pb;
  }
  if (threadIdx.x == 0) poff[V] = static_cast<uint32_t>(E);
  for (uint64_t k = threadIdx.x; k < E; k += blockDim.x) preds[k] = a.pred_rank[base + k] + 1;
  for (uint32_t c = threadIdx.x; c < a.stride + 16; c += blockDim.x) {
    a.seq_out[c] = c < a.n ? static_cast<uint8_t>(a.seq[c]) : 0u;
  }
}

// Turns the alignment kernel's result into the caller's record
// {n_pairs, score, has_score, status, width_bits} and reverses the pairs
// (stored in traceback order) into path order, in place.
__global__ void NoAllocFinish(const int32_t* res, int32_t* pairs, uint32_t cap, int32_t width_bits,
                              int32_t* result) {
  const int32_t status = res[3];
  const uint32_t np = static_cast<uint32_t>(res[0]);
  const bool ok = status == 0 && np <= cap;
  if (ok) {
    for (uint32_t k = threadIdx.x; k < np / 2; k += blockDim.x) {
      const int2 x = reinterpret_cast<int2*>(pairs)[k];
      const int2 y = reinterpret_cast<int2*>(pairs)[np - 1 - k];
      reinterpret_cast<int2*>(pairs)[k] = y;
      reinterpret_cast<int2*>(pairs)[np - 1 - k] = x;
    }
  }
  if (threadIdx.x == 0) {
    result[0] = ok ? static_cast<int32_t>(np) : 0;
    result[1] = res[1];
    result[2] = ok ? res[2] : 0;
    result[3] = status != 0 ? status : (np <= cap ? 0 : SPOA_HIP_CAPACITY);
    result[4] = width_bits;
  }
}

// The caller's result record when no DP runs (empty graph or read).
__global__ void NoAllocEmpty(int32_t* result) {
  if (threadIdx.x < 5) result[threadIdx.x] = 0;
}

}  // namespace spoa_hip_dev

namespace {

// Packs one job's graph block and read into the staging buffer.
// *r_out: R = sum over the rows of max(1, predecessors), for the DP's
// algorithmic bytes.
void PackJob(const spoa_hip_params& P, const spoa_hip_job& j, const Plan& pl, uint32_t* g, uint8_t* s,
             uint32_t* nslots_out, uint32_t* slot_chars_out, uint64_t* r_out) {
  const spoa_hip_graph& gr = j.graph;
  const uint32_t V = gr.num_nodes;
  uint32_t* rows = g;
  int32_t* node_id = reinterpret_cast<int32_t*>(g + 4 * V);
  uint32_t* poff = g + 5 * V;
  int32_t* fcH = reinterpret_cast<int32_t*>(g + 6 * V + 1);
  int32_t* fcF = fcH + (V + 1);
  int32_t* fcO = fcF + (V + 1);
  uint32_t* preds = g + 9 * V + 4;
  const uint32_t base = gr.pred_off[0];
  FirstColumns(P, gr, fcH, fcF, fcO);
  // profile slots: the distinct row bases, when there are at most 4
  int16_t slot_of[256];
  for (auto& x : slot_of) x = -1;
  uint32_t nslots = 0, slot_chars = 0;
  for (uint32_t r = 0; r < V && nslots <= 4; ++r) {
    const uint8_t ch = gr.row_char[r];
    if (slot_of[ch] < 0) {
      if (nslots < 4) slot_chars |= static_cast<uint32_t>(ch) << (8 * nslots);
      slot_of[ch] = static_cast<int16_t>(nslots++);
    }
  }
  if (nslots > 4) nslots = 0;
  *nslots_out = nslots;
  *slot_chars_out = slot_chars;
  uint64_t R = 0;
  for (uint32_t r = 0; r < V; ++r) {
    const uint32_t pb = gr.pred_off[r] - base, pe = gr.pred_off[r + 1] - base;
    const uint32_t npred = pe > pb ? pe - pb : 1u;
    R += npred;
    const uint32_t slot = nslots ? static_cast<uint32_t>(slot_of[gr.row_char[r]]) : 7u;
    node_id[r] = gr.node_id[r];
    poff[r] = pb;
    rows[4 * r + 0] = static_cast<uint32_t>(gr.row_char[r]) | (gr.is_sink[r] ? 0x100u : 0u) | (slot << 9) |
                      (npred << 12);
    rows[4 * r + 1] = pe > pb ? gr.pred_rank[gr.pred_off[r]] + 1 : 0u;
    rows[4 * r + 2] = pb;
    rows[4 * r + 3] = static_cast<uint32_t>(fcH[r + 1]);
  }
  poff[V] = gr.pred_off[V] - base;
  for (uint32_t k = base; k < gr.pred_off[V]; ++k) preds[k - base] = gr.pred_rank[k] + 1;
  *r_out = R;
  std::memcpy(s, j.seq, j.len);
  std::memset(s + j.len, 0, pl.seq_bytes - j.len);
}

// fn(a, b) over [0, n) in chunks: on the engine's host threads for large
// passes, on the calling thread for small ones.
template <typename Fn>
void ParallelRanges(spoa_hip_engine* E, size_t n, const Fn& fn) {
  constexpr size_t kChunk = 16;
  if (n < 4 * kChunk) {
    fn(size_t{0}, n);
    return;
  }
  if (!E->pool) E->pool.reset(new spoa_amd::ThreadPool(std::max(1u, std::min(16u, std::thread::hardware_concurrency()))));
  E->pool->For((n + kChunk - 1) / kChunk, [&](size_t c) { fn(c * kChunk, std::min(n, (c + 1) * kChunk)); });
}

// One pass: the jobs plans[lo, hi) (sorted by launch group, largest first
// inside a group) are packed into one staging block, copied in one
// transfer on the engine's copy stream, aligned by one persistent launch
// per group on the compute stream (one after another, each with the device
// to itself and timed by its own events) and copied back in one transfer.
// batch: a pass of the batched form (its arena starts at 4 GiB segments);
// else one per-call alignment (64 MiB, doubling as the graphs grow), so that
// an engine per host thread does not hold gigabytes for its first reads.
int RunPassLaunch(spoa_hip_engine* E, spoa_hip_job* jobs, std::vector<Plan>&& plans_in, size_t lo, size_t hi,
                  bool batch) {
  auto& pd = E->pending;
  pd.plans = std::move(plans_in);
  const std::vector<Plan>& plans = pd.plans;
  const spoa_hip_params& P = E->prm;
  const uint32_t nj = static_cast<uint32_t>(hi - lo);

  uint64_t graph_words = 0, seq_bytes = 0, pairs = 0;
  std::vector<uint64_t> g_at(nj), s_at(nj), p_at(nj);
  for (size_t x = lo; x < hi; ++x) {
    g_at[x - lo] = graph_words;
    s_at[x - lo] = seq_bytes;
    p_at[x - lo] = pairs;
    graph_words += plans[x].graph_words;
    seq_bytes += plans[x].seq_bytes;
    pairs += jobs[plans[x].job].graph.num_nodes + jobs[plans[x].job].len;
  }
  // split jobs: their hand-off words and the pass's queue items (every
  // tile of a split job is an item, the tiles of a job in order)
  std::vector<uint64_t> x_at(nj, 0);
  uint64_t xfer_words = spoa_hip_dev::kXferFirst, n_items = 0;
  bool any_split = false;
  for (size_t x = lo; x < hi; ++x) {
    n_items += plans[x].tiles;
    if (plans[x].tiles > 1) {
      any_split = true;
      x_at[x - lo] = xfer_words;
      xfer_words += XferWords(jobs[plans[x].job].graph.num_nodes, plans[x].tiles);
    }
  }
  if (xfer_words > 0xffffffffull) return Fail(SPOA_HIP_NO_MEMORY, E->msg->no_memory);
  const uint64_t off_jobs = 0;
  const uint64_t off_graph = RoundUp(nj * sizeof(DevJob), kAlign);
  const uint64_t off_seq = RoundUp(off_graph + graph_words * 4, kAlign);
  const uint64_t off_items = RoundUp(off_seq + seq_bytes, kAlign);
  const uint64_t in_bytes = RoundUp(off_items + (any_split ? n_items * sizeof(uint2) : 0), kAlign);
  const uint64_t off_res = 0;
  const uint64_t off_pairs = RoundUp(static_cast<uint64_t>(nj) * 16, kAlign);
  const uint64_t out_bytes = off_pairs + pairs * 8;

  const auto t_alloc0 = std::chrono::steady_clock::now();
  if (!E->h_in.Reserve(in_bytes) || !E->h_out.Reserve(out_bytes)) return SPOA_HIP_DEVICE_ERROR;
  if (!E->d_in.Reserve(in_bytes, E->msg->no_memory) || !E->d_out.Reserve(out_bytes, E->msg->no_memory) ||
      !E->d_counter.Reserve(256 * sizeof(uint32_t), E->msg->no_memory)) {
    return SPOA_HIP_NO_MEMORY;
  }
  if (any_split && !ReserveXfer(E, xfer_words)) return SPOA_HIP_NO_MEMORY;
  // place every job's matrices in an arena segment (first fit in order)
  std::vector<uint64_t> addr(nj);
  {
    std::vector<uint64_t> need(nj);
    for (size_t x = 0; x < nj; ++x) need[x] = plans[lo + x].mat_bytes;
    const int rc = PlaceMatrices(E, need, &addr, batch ? kBatchSegment : kPerCallSegment);
    if (rc != SPOA_HIP_OK) return rc;
  }
  E->stats.alloc_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_alloc0).count();

  // ---- pack (host threads for large passes)
  const auto t_host0 = std::chrono::steady_clock::now();
  DevJob* dj = E->h_in.At<DevJob>(off_jobs);
  std::vector<uint64_t> job_r(nj);
  auto pack_range = [&](size_t a, size_t b) {
    for (size_t x = a; x < b; ++x) {
      const Plan& pl = plans[lo + x];
      const spoa_hip_job& j = jobs[pl.job];
      DevJob& d = dj[x];
      d.mat_addr = addr[x];
      d.graph_off = static_cast<uint32_t>(g_at[x]);
      d.seq_off = static_cast<uint32_t>(s_at[x]);
      d.V = j.graph.num_nodes;
      d.n = j.len;
      d.W = pl.W;
      d.stride = pl.stride;
      d.pair_off = static_cast<uint32_t>(p_at[x]);
      d.cap = j.graph.num_nodes + j.len;
      d.split = pl.tiles > 1 ? pl.tiles : 0u;
      d.xfer_off = static_cast<uint32_t>(x_at[x]);
      d.pad0 = d.pad1 = 0;
      PackJob(P, j, pl, E->h_in.At<uint32_t>(off_graph) + g_at[x], E->h_in.At<uint8_t>(off_seq) + s_at[x],
              &d.nslots, &d.slot_chars, &job_r[x]);
    }
  };
  // Passes of 128 MiB or more on the stream path are packed in slices of
  // jobs, each slice's three regions (job records, graph blocks, reads)
  // copied in as soon as it is packed, so most of the transfer runs while
  // the later slices are packed.  Smaller passes (and captured ones) copy
  // everything in one transfer: there the extra copies cost more host time
  // than the overlap saves (measured on passes of 17 and 86 MB).
  const size_t nslices = !E->graph_mode && nj >= 4 && in_bytes >= (128ull << 20) ? 4 : 1;
  if (nslices > 1 && !HipOk(hipEventRecord(E->ev0, E->copy), "event")) return SPOA_HIP_DEVICE_ERROR;
  for (size_t sl = 0; sl < nslices; ++sl) {
    const size_t a = nj * sl / nslices, b = nj * (sl + 1) / nslices;
    ParallelRanges(E, b - a, [&](size_t u, size_t v) { pack_range(a + u, a + v); });
    if (nslices == 1) break;
    const uint64_t g0 = g_at[a] * 4, g1 = (b < nj ? g_at[b] : graph_words) * 4;
    const uint64_t s0 = s_at[a], s1 = b < nj ? s_at[b] : seq_bytes;
    uint8_t* din0 = static_cast<uint8_t*>(E->d_in.p);
    const uint8_t* hin0 = E->h_in.At<uint8_t>(0);
    auto copy = [&](uint64_t off, uint64_t bytes) {
      return bytes == 0 ||
             HipOk(hipMemcpyAsync(din0 + off, hin0 + off, bytes, hipMemcpyHostToDevice, E->copy), "copy in");
    };
    if (!copy(off_jobs + a * sizeof(DevJob), (b - a) * sizeof(DevJob)) || !copy(off_graph + g0, g1 - g0) ||
        !copy(off_seq + s0, s1 - s0)) {
      return SPOA_HIP_DEVICE_ERROR;
    }
  }
  if (any_split) {
    // the item table: (job index inside its launch group, tile)
    uint2* it = E->h_in.At<uint2>(off_items);
    size_t g0 = lo;
    for (size_t x = lo; x < hi; ++x) {
      if (x > lo && !SameGroup(P.subtype, plans[x], plans[x - 1])) g0 = x;
      for (uint32_t t = 0; t < plans[x].tiles; ++t) *it++ = make_uint2(static_cast<uint32_t>(x - g0), t);
    }
    if (nslices > 1 && !HipOk(hipMemcpyAsync(static_cast<uint8_t*>(E->d_in.p) + off_items, E->h_in.At<uint8_t>(off_items),
                                             n_items * sizeof(uint2), hipMemcpyHostToDevice, E->copy),
                              "copy in")) {
      return SPOA_HIP_DEVICE_ERROR;
    }
  }
  // algorithmic DP bytes of the pass's jobs: S * n * (M * V + Rd * R)
  for (size_t x = 0; x < nj; ++x) {
    const Plan& pl = plans[lo + x];
    const spoa_hip_job& j = jobs[pl.job];
    const uint64_t jb = static_cast<uint64_t>(pl.bits / 8) * j.len *
                        (static_cast<uint64_t>(Matrices(P.subtype)) * j.graph.num_nodes +
                         static_cast<uint64_t>(ReadMatrices(P.subtype)) * job_r[x]);
    E->stats.dp_bytes += jb;
    (pl.width == 16 ? E->stats.dp_bytes16 : E->stats.dp_bytes32) += jb;
  }
  E->stats.host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host0).count();

  // ---- launch groups
  std::vector<LaunchGroup> groups;
  {
    const int rc = BuildGroups(E, plans, lo, hi, &groups);
    if (rc != SPOA_HIP_OK) return rc;
  }
  if (groups.size() > 256) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] too many launch groups");

  while (E->ev_ks.size() < groups.size()) {
    hipEvent_t a = nullptr, b = nullptr;
    if (!HipOk(hipEventCreate(&a), "hipEventCreate")) return SPOA_HIP_DEVICE_ERROR;
    if (!HipOk(hipEventCreate(&b), "hipEventCreate")) {
      (void)hipEventDestroy(a);
      return SPOA_HIP_DEVICE_ERROR;
    }
    E->ev_ks.push_back(a);
    E->ev_ke.push_back(b);
  }

  uint8_t* din = static_cast<uint8_t*>(E->d_in.p);
  uint8_t* dout = static_cast<uint8_t*>(E->d_out.p);
  uint32_t* counters = static_cast<uint32_t*>(E->d_counter.p);
  // every launch tags its hand-off words with an epoch of its own
  std::vector<uint32_t> epochs(groups.size());
  for (auto& ep : epochs) ep = NextEpoch(E);
  hipStream_t st = E->copy;
  hipStream_t cs = E->compute;
  auto args_of = [&](size_t gi) {
    const LaunchGroup& gp = groups[gi];
    KArgs args{};
    args.prm = gp.prm;
    args.jobs = reinterpret_cast<const DevJob*>(din + off_jobs) + gp.a;
    args.n_jobs = static_cast<uint32_t>(gp.b - gp.a);
    args.queue = counters + gi;
    args.graphs = reinterpret_cast<const uint32_t*>(din + off_graph);
    args.seqs = din + off_seq;
    args.pairs = reinterpret_cast<int32_t*>(dout + off_pairs);
    args.results = reinterpret_cast<int32_t*>(dout + off_res) + 4 * gp.a;
    args.lds_bytes = gp.lds;
    args.prio_jobs = gp.prio;
    args.phases = E->phases;
    if (any_split) {
      args.items = reinterpret_cast<const uint2*>(din + off_items) + gp.item_at;
      args.n_items = gp.items;
    }
    args.epoch = epochs[gi];
    args.wait_ticks = E->wait_ticks;
    args.xfer = static_cast<unsigned long long*>(E->d_xfer.p);
    return args;
  };
  E->pass_width.clear();
  // statistics by the reference's width (its int32 jobs on int16 rows count
  // with the int32 launches, whose launch shape they take)
  for (const LaunchGroup& gp : groups) E->pass_width.push_back(plans[lo + gp.a].width);
  E->stats.launches += static_cast<uint32_t>(groups.size());
  E->graph_pass = E->graph_mode;
  if (E->graph_mode) {
    // capture copy in, counter reset (a kernel, not a memset node), the
    // group launches and copy out on the copy stream, then launch the graph
    // on the compute stream; its device time is timed as one unit
    if (!HipOk(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "graph capture")) {
      return SPOA_HIP_DEVICE_ERROR;
    }
    bool ok = HipOk(hipMemcpyAsync(E->d_in.p, E->h_in.p, in_bytes, hipMemcpyHostToDevice, st), "copy in");
    if (ok) {
      spoa_hip_dev::ResetCounters<<<1, 256, 0, st>>>(counters, static_cast<uint32_t>(groups.size()));
      ok = HipOk(hipGetLastError(), "counter reset");
    }
    for (size_t gi = 0; ok && gi < groups.size(); ++gi) {
      const LaunchGroup& gp = groups[gi];
      gp.fn<<<gp.grid, gp.threads, gp.lds, st>>>(args_of(gi));
      ok = HipOk(hipGetLastError(), "kernel launch");
    }
    if (ok) ok = HipOk(hipMemcpyAsync(E->h_out.p, E->d_out.p, out_bytes, hipMemcpyDeviceToHost, st), "copy out");
    hipGraph_t gr = nullptr;
    const bool ended = HipOk(hipStreamEndCapture(st, &gr), "graph capture");
    if (!ok || !ended) {
      if (gr) (void)hipGraphDestroy(gr);
      return SPOA_HIP_DEVICE_ERROR;
    }
    // The previous pass's executable graph takes this pass's nodes in place
    // (same kernels and copies; new grid sizes, arguments and byte counts)
    // when the two captures have the same shape; otherwise the capture is
    // instantiated anew.  The previous launch of it has finished (its pass
    // was waited for before this one was built).
    bool updated = false;
    if (E->graph_exec) {
      hipGraphNode_t err_node = nullptr;
      hipGraphExecUpdateResult res = hipGraphExecUpdateError;
      if (hipGraphExecUpdate(E->graph_exec, gr, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
        updated = true;
      } else {
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(E->graph_exec);
        E->graph_exec = nullptr;
      }
    }
    const bool built = updated || HipOk(hipGraphInstantiate(&E->graph_exec, gr, nullptr, nullptr, 0), "graph instantiate");
    (void)hipGraphDestroy(gr);
    if (built) ++(updated ? E->stats.graph_updates : E->stats.graph_builds);
    if (!built || !HipOk(hipEventRecord(E->ev_ks[0], cs), "event") || !HipOk(hipGraphLaunch(E->graph_exec, cs), "graph launch") ||
        !HipOk(hipEventRecord(E->ev_ke[0], cs), "event") || !HipOk(hipEventRecord(E->ev_kdone, cs), "event")) {
      return SPOA_HIP_DEVICE_ERROR;
    }
    E->pass_launches = 1;
    E->pass_width.assign(1, plans[lo].width);
  } else {
    if (nslices == 1 &&
        (!HipOk(hipEventRecord(E->ev0, st), "event") ||
         !HipOk(hipMemcpyAsync(E->d_in.p, E->h_in.p, in_bytes, hipMemcpyHostToDevice, st), "copy in"))) {
      return SPOA_HIP_DEVICE_ERROR;
    }
    if (!HipOk(hipEventRecord(E->ev1, st), "event")) return SPOA_HIP_DEVICE_ERROR;
    if (cs != st && (!HipOk(hipEventRecord(E->ev_in, st), "event") ||
                     !HipOk(hipStreamWaitEvent(cs, E->ev_in, 0), "stream wait"))) {
      return SPOA_HIP_DEVICE_ERROR;
    }
    // The first group runs on the compute stream, the others beside it on
    // the side stream (one after another there), each launch timed by its
    // own events; the copy out waits for both streams.  Each stream zeroes
    // the job counters of its own launches, so the side launches wait only
    // for the copy in, not for the compute stream's earlier work.
    if (groups.size() > 1 && !E->side) {
      if (!E->side_own && !HipOk(hipStreamCreateWithFlags(&E->side_own, hipStreamNonBlocking), "hipStreamCreate")) {
        return SPOA_HIP_DEVICE_ERROR;
      }
      E->side = E->side_own;
    }
    const bool beside = groups.size() > 1 && E->side != cs;
    spoa_hip_dev::ResetCounters<<<1, 256, 0, cs>>>(counters, beside ? 1u : static_cast<uint32_t>(groups.size()));
    if (!HipOk(hipGetLastError(), "counter reset")) return SPOA_HIP_DEVICE_ERROR;
    if (beside) {
      if (E->side != st && (!HipOk(hipEventRecord(E->ev_side_in, st), "event") ||
                            !HipOk(hipStreamWaitEvent(E->side, E->ev_side_in, 0), "stream wait"))) {
        return SPOA_HIP_DEVICE_ERROR;
      }
      spoa_hip_dev::ResetCounters<<<1, 256, 0, E->side>>>(counters + 1, static_cast<uint32_t>(groups.size() - 1));
      if (!HipOk(hipGetLastError(), "counter reset")) return SPOA_HIP_DEVICE_ERROR;
    }
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      const LaunchGroup& gp = groups[gi];
      hipStream_t ks = gi > 0 && beside ? E->side : cs;
      if (!HipOk(hipEventRecord(E->ev_ks[gi], ks), "event")) return SPOA_HIP_DEVICE_ERROR;
      gp.fn<<<gp.grid, gp.threads, gp.lds, ks>>>(args_of(gi));
      if (!HipOk(hipGetLastError(), "kernel launch") || !HipOk(hipEventRecord(E->ev_ke[gi], ks), "event")) {
        return SPOA_HIP_DEVICE_ERROR;
      }
    }
    E->pass_launches = groups.size();
    // the copy out waits for both streams; the compute stream itself does
    // not wait for the side stream, so the next pass's first launch (on a
    // shared compute stream) can run beside this pass's side launches
    if (cs != st && (!HipOk(hipEventRecord(E->ev_kdone, cs), "event") ||
                     !HipOk(hipStreamWaitEvent(st, E->ev_kdone, 0), "stream wait"))) {
      return SPOA_HIP_DEVICE_ERROR;
    }
    if (beside && (!HipOk(hipEventRecord(E->ev_side, E->side), "event") ||
                   !HipOk(hipStreamWaitEvent(st, E->ev_side, 0), "stream wait"))) {
      return SPOA_HIP_DEVICE_ERROR;
    }
    if (!HipOk(hipEventRecord(E->ev2, st), "event") ||
        !HipOk(hipMemcpyAsync(E->h_out.p, E->d_out.p, out_bytes, hipMemcpyDeviceToHost, st), "copy out") ||
        !HipOk(hipEventRecord(E->ev3, st), "event")) {
      return SPOA_HIP_DEVICE_ERROR;
    }
  }
  pd.active = true;
  pd.jobs = jobs;
  pd.lo = lo;
  pd.hi = hi;
  pd.off_res = off_res;
  pd.off_pairs = off_pairs;
  pd.pair_at = std::move(p_at);
  return SPOA_HIP_OK;
}

// Waits for the pending pass and unpacks its results into the job records.
int RunPassFinish(spoa_hip_engine* E) {
  auto& pd = E->pending;
  if (!pd.active) return SPOA_HIP_OK;
  pd.active = false;
  spoa_hip_job* jobs = pd.jobs;
  const std::vector<Plan>& plans = pd.plans;
  const size_t lo = pd.lo, hi = pd.hi;
  hipStream_t st = E->copy;
  const auto t_wait0 = std::chrono::steady_clock::now();
  if (E->graph_pass) {
    // the executable graph stays for the next pass to update
    if (!HipOk(hipEventSynchronize(E->ev_kdone), "graph")) return SPOA_HIP_DEVICE_ERROR;
  } else if (!HipOk(hipStreamSynchronize(st), "kernel")) {
    return SPOA_HIP_DEVICE_ERROR;
  }
  const auto t_wait1 = std::chrono::steady_clock::now();
  E->stats.wait_ms += std::chrono::duration<double, std::milli>(t_wait1 - t_wait0).count();
  if (!E->graph_pass) {
    float ms_in = 0, ms_out = 0;
    (void)hipEventElapsedTime(&ms_in, E->ev0, E->ev1);
    (void)hipEventElapsedTime(&ms_out, E->ev2, E->ev3);
    E->stats.h2d_ms += ms_in;
    E->stats.d2h_ms += ms_out;
  }
  for (size_t gi = 0; gi < E->pass_launches; ++gi) {
    float ms_k = 0;
    (void)hipEventElapsedTime(&ms_k, E->ev_ks[gi], E->ev_ke[gi]);
    E->stats.kernel_ms += ms_k;
    if (gi < E->pass_width.size()) {
      if (E->pass_width[gi] == 16) {
        E->stats.kernel_ms16 += ms_k;
        E->stats.launches16 += 1;
      } else {
        E->stats.kernel_ms32 += ms_k;
        E->stats.launches32 += 1;
      }
    }
  }

  // ---- unpack: pairs come back in traceback order
  const int32_t* res = E->h_out.At<int32_t>(pd.off_res);
  const int32_t* allp = E->h_out.At<int32_t>(pd.off_pairs);
  auto unpack_range = [&](size_t a, size_t b) {
    for (size_t x = lo + a; x < lo + b; ++x) {
      spoa_hip_job& j = jobs[plans[x].job];
      const int32_t* r4 = res + 4 * (x - lo);
      j.width_bits = plans[x].width;
      j.has_score = r4[2];
      j.score = r4[1];
      const uint32_t np = static_cast<uint32_t>(r4[0]);
      int status = r4[3] == spoa_hip_dev::kDevCapacity       ? SPOA_HIP_CAPACITY
                   : r4[3] == spoa_hip_dev::kDevInconsistent ? SPOA_HIP_INCONSISTENT
                   : r4[3] == spoa_hip_dev::kDevTimeout      ? SPOA_HIP_DEVICE_ERROR
                                                             : SPOA_HIP_OK;
      if (status == SPOA_HIP_OK && np > j.cap) status = SPOA_HIP_CAPACITY;
      if (status == SPOA_HIP_OK) {
        const int32_t* src = allp + 2ull * pd.pair_at[x - lo];
        for (uint32_t k = 0; k < np; ++k) {
          j.pairs[2 * k] = src[2 * (np - 1 - k)];
          j.pairs[2 * k + 1] = src[2 * (np - 1 - k) + 1];
        }
        j.n_pairs = np;
      } else {
        j.n_pairs = 0;
      }
      j.status = status;
    }
  };
  ParallelRanges(E, hi - lo, unpack_range);
  // the first failed job (in pass order) names the error
  int first = SPOA_HIP_OK;
  for (size_t x = lo; x < hi && first == SPOA_HIP_OK; ++x) {
    const int status = jobs[plans[x].job].status;
    if (status != SPOA_HIP_OK) {
      first = status;
      g_last_error = status == SPOA_HIP_CAPACITY       ? "[spoa::hip] alignment output capacity exceeded"
                     : status == SPOA_HIP_DEVICE_ERROR ? kSplitTimeoutMsg
                                                       : "[spoa::hip] traceback found no consistent move";
    }
  }
  E->stats.unpack_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wait1).count();
  return first;
}

spoa_amd::PassItem PassItemOf(const Plan& pl, const spoa_hip_job& j) {
  return spoa_amd::PassItem{pl.mat_bytes, pl.graph_words, pl.seq_bytes,
                            static_cast<uint64_t>(j.graph.num_nodes) + j.len};
}

int RunBatch(spoa_hip_engine* E, spoa_hip_job* jobs, uint32_t n_jobs, bool batch, bool async) {
  if (E->pending.active) {
    const int rc = RunPassFinish(E);
    (void)rc;
  }
  E->stats = spoa_hip_stats{};
  E->pending.first = SPOA_HIP_OK;
  const spoa_hip_params& P = E->prm;
  int first = SPOA_HIP_OK;
  auto fail = [&](spoa_hip_job& j, int code, const char* msg) {
    j.status = code;
    j.n_pairs = 0;
    j.has_score = 0;
    if (first == SPOA_HIP_OK) {
      first = code;
      g_last_error = msg;
    }
  };
  // A batch runs one launch configuration per score width, the one for its
  // longest read at that width: jobs of one width then share one launch
  // (shorter rows leave lanes idle) instead of several launches in turn.
  uint32_t cfg_len[2] = {0, 0};
  if (batch) {
    for (uint32_t x = 0; x < n_jobs; ++x) {
      const spoa_hip_job& j = jobs[x];
      if (j.graph.num_nodes == 0 || j.len == 0 || j.len > static_cast<uint32_t>(std::numeric_limits<int32_t>::max())) {
        continue;
      }
      const Shape sh = JobShape(P, j.len, j.graph.num_nodes, E->sisd);
      if (sh.width != 0) cfg_len[sh.cfgw == 32] = std::max(cfg_len[sh.cfgw == 32], j.len);
    }
  }
  std::vector<Plan> plans;
  plans.reserve(n_jobs);
  for (uint32_t x = 0; x < n_jobs; ++x) {
    spoa_hip_job& j = jobs[x];
    j.n_pairs = 0;
    j.has_score = 0;
    j.width_bits = 0;
    j.status = SPOA_HIP_OK;
    // reference simd...:708-715
    if (j.len > static_cast<uint32_t>(std::numeric_limits<int32_t>::max())) {
      fail(j, SPOA_HIP_TOO_LARGE, E->msg->too_large);
      continue;
    }
    const uint32_t V = j.graph.num_nodes;
    if (V == 0 || j.len == 0) continue;
    const Shape sh = JobShape(P, j.len, V, E->sisd);
    if (sh.width == 0) {
      fail(j, SPOA_HIP_OVERFLOW, E->msg->overflow);
      continue;
    }
    Plan pl;
    pl.job = x;
    pl.width = sh.width;
    pl.bits = sh.bits;
    pl.lanes = sh.lanes;
    // launch configuration of the reference width (int16 rows of the
    // reference's int32 jobs take the int32 launch shape: they form the
    // same few, long launches)
    pl.cfg = ChooseCfg(E, batch ? cfg_len[sh.cfgw == 32] : j.len, sh.cfgw, batch);
    pl.work = static_cast<uint64_t>(V) * j.len;
    const int L = sh.lanes;
    pl.W = (j.len + L - 1) / L * L;
    pl.stride = static_cast<uint32_t>(RoundUp(pl.W, static_cast<uint64_t>(pl.cfg.k)));
    pl.tiles = JobTiles(pl.stride, pl.cfg);
    if (pl.tiles > kMaxSplitTiles) {  // such long rows run their tiles in turn
      pl.cfg.split = 0;
      pl.tiles = 1;
    }
    // a one-tile row is not split: the job joins the launches whose kernel
    // carries no hand-off code
    if (pl.tiles == 1) pl.cfg.split = 0;
    pl.mat_bytes = JobBytes(P, V, pl.stride, sh.bits, pl.cfg);
    const uint32_t E_in = j.graph.pred_off[V] - j.graph.pred_off[0];
    pl.graph_words = RoundUp(9ull * V + 4 + E_in, 4);
    pl.seq_bytes = RoundUp(pl.stride + 16, 16);
    if (!spoa_amd::PassItemFits(PassItemOf(pl, j), E->budget)) {
      fail(j, SPOA_HIP_NO_MEMORY, E->msg->no_memory);
      continue;
    }
    plans.push_back(pl);
    // algorithmic work of this DP (its bytes are counted as the job is packed)
    E->stats.cells += static_cast<uint64_t>(V) * j.len;
    E->stats.jobs += 1;
  }
  // group by (row width, lanes, cfg); within a group the queue runs largest V * n first
  std::stable_sort(plans.begin(), plans.end(), [&](const Plan& a, const Plan& b) {
    if (a.bits != b.bits) return a.bits < b.bits;
    const int la = GroupLanes(P.subtype, a), lb = GroupLanes(P.subtype, b);
    if (la != lb) return la > lb;
    if (a.cfg.k != b.cfg.k) return a.cfg.k < b.cfg.k;
    if (a.cfg.waves != b.cfg.waves) return a.cfg.waves < b.cfg.waves;
    if (a.cfg.split != b.cfg.split) return a.cfg.split < b.cfg.split;
    const uint64_t wa = static_cast<uint64_t>(jobs[a.job].graph.num_nodes) * jobs[a.job].len;
    const uint64_t wb = static_cast<uint64_t>(jobs[b.job].graph.num_nodes) * jobs[b.job].len;
    return wa > wb;
  });
  // passes: DP matrices within the memory budget, 32-bit offsets in range
  std::vector<spoa_amd::PassItem> items(plans.size());
  for (size_t x = 0; x < plans.size(); ++x) items[x] = PassItemOf(plans[x], jobs[plans[x].job]);
  size_t lo = 0;
  while (lo < plans.size()) {
    const size_t hi = spoa_amd::PassEnd(items.data(), lo, items.size(), E->budget);
    const bool last = hi == plans.size();
    // the pass owns a copy of the plans while it is in flight
    int rc = RunPassLaunch(E, jobs, std::vector<Plan>(plans), lo, hi, batch);
    if (rc == SPOA_HIP_OK && !(last && async)) rc = RunPassFinish(E);
    if (rc != SPOA_HIP_OK && first == SPOA_HIP_OK) first = rc;
    if (rc == SPOA_HIP_DEVICE_ERROR || rc == SPOA_HIP_NO_MEMORY || rc == SPOA_HIP_INVALID_ARGUMENT) {
      const std::string msg = g_last_error;
      E->pending.active = false;
      for (size_t x = lo; x < plans.size(); ++x) jobs[plans[x].job].status = rc;
      g_last_error = msg;
      return rc;
    }
    lo = hi;
  }
  E->pending.first = first;
  return first;
}

}  // namespace

extern "C" {

const char* spoa_hip_last_error(void) { return g_last_error.c_str(); }

int spoa_hip_engine_create(int device, const spoa_hip_params* params, spoa_hip_engine** out) {
  if (!params || !out) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null argument");
  if (params->type < 0 || params->type > 2) {
    return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::AlignmentEngine::Create] error: invalid alignment type!");
  }
  if (params->subtype < 0 || params->subtype > 2) {
    return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] invalid alignment subtype");
  }
  int count = 0;
  if (!HipOk(hipGetDeviceCount(&count), "hipGetDeviceCount")) return SPOA_HIP_DEVICE_ERROR;
  if (device < 0 || device >= count) return Fail(SPOA_HIP_DEVICE_ERROR, "[spoa::hip] no such HIP device");
  if (!HipOk(hipSetDevice(device), "hipSetDevice")) return SPOA_HIP_DEVICE_ERROR;
  if (ScratchInit(device) != SPOA_HIP_OK) return SPOA_HIP_DEVICE_ERROR;
  auto* E = new spoa_hip_engine();
  E->device = device;
  E->prm = *params;
  hipDeviceProp_t prop;
  if (HipOk(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties")) E->num_cus = prop.multiProcessorCount;
  size_t free_b = 0, total_b = 0;
  if (HipOk(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo")) {
    E->budget = static_cast<uint64_t>(static_cast<double>(free_b) * 0.80);
  } else {
    E->budget = 8ull << 30;
  }
  if (!HipOk(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking), "hipStreamCreate") ||
      !HipOk(hipEventCreate(&E->ev0), "hipEventCreate") || !HipOk(hipEventCreate(&E->ev1), "hipEventCreate") ||
      !HipOk(hipEventCreate(&E->ev2), "hipEventCreate") || !HipOk(hipEventCreate(&E->ev3), "hipEventCreate") ||
      !HipOk(hipEventCreateWithFlags(&E->ev_in, hipEventDisableTiming), "hipEventCreate") ||
      !HipOk(hipEventCreateWithFlags(&E->ev_kdone, hipEventDisableTiming), "hipEventCreate") ||
      !HipOk(hipEventCreateWithFlags(&E->ev_side_in, hipEventDisableTiming), "hipEventCreate") ||
      !HipOk(hipEventCreateWithFlags(&E->ev_side, hipEventDisableTiming), "hipEventCreate")) {
    spoa_hip_engine_destroy(E);
    return SPOA_HIP_DEVICE_ERROR;
  }
  E->compute = E->stream;
  E->copy = E->stream;
#ifdef SPOA_ROW_PHASES
  if (!HipOk(hipMalloc(&E->phases, 64 * sizeof(unsigned long long)), "hipMalloc") ||
      !HipOk(hipMemset(E->phases, 0, 64 * sizeof(unsigned long long)), "hipMemset")) {
    spoa_hip_engine_destroy(E);
    return SPOA_HIP_DEVICE_ERROR;
  }
#endif
  *out = E;
  return SPOA_HIP_OK;
}

int spoa_hip_row_phases(spoa_hip_engine* E, uint64_t* out, uint32_t n, int32_t reset) {
  if (!E || (!out && n)) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null argument");
  if (!E->phases) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] built without row-phase timing");
  if (!HipOk(hipSetDevice(E->device), "hipSetDevice") || !HipOk(hipDeviceSynchronize(), "hipDeviceSynchronize")) {
    return SPOA_HIP_DEVICE_ERROR;
  }
  const uint32_t k = std::min<uint32_t>(n, spoa_hip_dev::kPhTotal);
  if (k && !HipOk(hipMemcpy(out, E->phases, k * sizeof(uint64_t), hipMemcpyDeviceToHost), "copy out")) {
    return SPOA_HIP_DEVICE_ERROR;
  }
  if (reset && !HipOk(hipMemset(E->phases, 0, 64 * sizeof(unsigned long long)), "hipMemset")) {
    return SPOA_HIP_DEVICE_ERROR;
  }
  return static_cast<int>(k);
}

int spoa_hip_engine_create_sisd(int device, const spoa_hip_params* params, spoa_hip_engine** out) {
  const int rc = spoa_hip_engine_create(device, params, out);
  if (rc != SPOA_HIP_OK) return rc;
  (*out)->sisd = true;
  (*out)->msg = &kSisdMsgs;
  return SPOA_HIP_OK;
}

int spoa_hip_engine_destroy(spoa_hip_engine* E) {
  if (!E) return SPOA_HIP_OK;
  (void)hipSetDevice(E->device);
  if (E->pending.active) (void)RunPassFinish(E);
  if (E->stream) (void)hipStreamSynchronize(E->stream);
  ReleaseWinBuffers(E);
  E->d_in.Release();
  E->d_out.Release();
  for (auto& sg : E->arena) sg.Release();
  E->arena.clear();
  E->d_counter.Release();
  E->h_in.Release();
  E->h_out.Release();
  if (E->ev0) (void)hipEventDestroy(E->ev0);
  if (E->ev1) (void)hipEventDestroy(E->ev1);
  if (E->ev2) (void)hipEventDestroy(E->ev2);
  if (E->ev3) (void)hipEventDestroy(E->ev3);
  if (E->graph_exec) (void)hipGraphExecDestroy(E->graph_exec);
  if (E->ev_in) (void)hipEventDestroy(E->ev_in);
  if (E->ev_kdone) (void)hipEventDestroy(E->ev_kdone);
  if (E->ev_side_in) (void)hipEventDestroy(E->ev_side_in);
  if (E->ev_side) (void)hipEventDestroy(E->ev_side);
  if (E->side_own) {
    (void)hipStreamSynchronize(E->side_own);
    (void)hipStreamDestroy(E->side_own);
  }
  for (auto ev : E->ev_ks) (void)hipEventDestroy(ev);
  for (auto ev : E->ev_ke) (void)hipEventDestroy(ev);
  if (E->compute_own) {
    (void)hipStreamSynchronize(E->compute_own);
    (void)hipStreamDestroy(E->compute_own);
  }
  if (E->stream) (void)hipStreamDestroy(E->stream);
  if (E->phases) (void)hipFree(E->phases);
  delete E;
  return SPOA_HIP_OK;
}

// reference simd...:381-422; the scalar engine (sisd...:62-77) checks only
// the length and sizes int32 rows with no overflow bound
int spoa_hip_prealloc(spoa_hip_engine* E, uint32_t max_len, uint32_t alphabet_size) {
  if (!E) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null engine");
  if (max_len > static_cast<uint32_t>(std::numeric_limits<int32_t>::max())) {
    return Fail(SPOA_HIP_TOO_LARGE, E->msg->pre_too_large);
  }
  const int64_t V = static_cast<int64_t>(max_len) * (alphabet_size & 0xffu);
  int width = 32;
  if (!E->sisd) {
    const int64_t w = WorstCase(E->prm, static_cast<int64_t>(max_len) + 8, V);
    if (w < std::numeric_limits<int32_t>::min() + 1024) return SPOA_HIP_OK;
    width = w < std::numeric_limits<int16_t>::min() + 1024 ? 32 : 16;
  }
  if (!HipOk(hipSetDevice(E->device), "hipSetDevice")) return SPOA_HIP_DEVICE_ERROR;
  const Cfg cfg = ChooseCfg(E, max_len, width, false);
  const int L = E->sisd ? 1 : Lanes(width);  // the reference's sizing (int32 rows at most)
  const uint64_t W = (static_cast<uint64_t>(max_len) + L - 1) / L * L;
  const uint64_t stride = RoundUp(W, static_cast<uint64_t>(cfg.k));
  if (V >= static_cast<int64_t>(std::numeric_limits<uint32_t>::max())) return Fail(SPOA_HIP_NO_MEMORY, E->msg->pre_no_memory);
  const uint64_t bytes = JobBytes(E->prm, static_cast<uint32_t>(V), static_cast<uint32_t>(stride), width, cfg);
  if (bytes > E->budget) return Fail(SPOA_HIP_NO_MEMORY, E->msg->pre_no_memory);
  uint64_t largest = 0;
  for (const auto& sg : E->arena) largest = std::max(largest, sg.size);
  if (largest < bytes) {
    // the arena stays within the budget: when the new segment would take it
    // past, the older (smaller) segments go first -- the passes that follow
    // grow the arena again within the budget as they need
    if (E->arena_bytes + bytes > E->budget) {
      if (E->pending.active) (void)RunPassFinish(E);
      if (!HipOk(hipStreamSynchronize(E->stream), "hipStreamSynchronize")) return SPOA_HIP_DEVICE_ERROR;
      for (auto& sg : E->arena) sg.Release();
      E->arena.clear();
      E->arena_bytes = 0;
    }
    DevBuf sg;
    if (hipMalloc(&sg.p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return Fail(SPOA_HIP_NO_MEMORY, E->msg->pre_no_memory);
    }
    sg.size = bytes;
    E->arena.push_back(sg);
    E->arena_bytes += bytes;
  }
  return SPOA_HIP_OK;
}

int spoa_hip_align(spoa_hip_engine* E, const spoa_hip_graph* graph, const char* seq, uint32_t len, int32_t* pairs,
                   uint32_t cap, uint32_t* n_pairs, int32_t* score, int32_t* has_score) {
  if (!E || !graph || (!seq && len) || !n_pairs) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null argument");
  if (!HipOk(hipSetDevice(E->device), "hipSetDevice")) return SPOA_HIP_DEVICE_ERROR;
  spoa_hip_job j{};
  j.graph = *graph;
  j.seq = seq;
  j.len = len;
  j.pairs = pairs;
  j.cap = cap;
  const int rc = RunBatch(E, &j, 1, false, false);
  *n_pairs = j.n_pairs;
  if (has_score) *has_score = j.has_score;
  if (score && j.has_score) *score = j.score;
  return rc;
}

int spoa_hip_align_noalloc(spoa_hip_engine* E, const spoa_hip_graph* graph, const char* seq, uint32_t len,
                           int32_t* pairs, uint32_t cap, int32_t* result, void* stream) {
  if (!E || !graph || !result || (!seq && len) || (!pairs && cap)) {
    return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null argument");
  }
  if (!HipOk(hipSetDevice(E->device), "hipSetDevice")) return SPOA_HIP_DEVICE_ERROR;
  uint8_t* const base = E->device < kMaxDevices ? g_scratch[E->device] : nullptr;
  if (!base) return Fail(SPOA_HIP_DEVICE_ERROR, "[spoa::hip] no scratch address for the device");
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : E->stream;
  // reference simd...:708-715
  if (len > static_cast<uint32_t>(std::numeric_limits<int32_t>::max())) return Fail(SPOA_HIP_TOO_LARGE, E->msg->too_large);
  const uint32_t V = graph->num_nodes;
  if (V == 0 || len == 0) {
    spoa_hip_dev::NoAllocEmpty<<<1, 64, 0, st>>>(result);
    return HipOk(hipGetLastError(), "kernel launch") ? SPOA_HIP_OK : SPOA_HIP_DEVICE_ERROR;
  }
  const Shape sh = JobShape(E->prm, len, V, E->sisd);
  if (sh.width == 0) return Fail(SPOA_HIP_OVERFLOW, E->msg->overflow);
  // the per-call launch configuration of spoa_hip_align
  const Cfg cfg = ChooseCfg(E, len, sh.cfgw, false);
  const uint32_t W = (len + sh.lanes - 1) / sh.lanes * sh.lanes;
  const uint64_t stride = RoundUp(W, static_cast<uint64_t>(cfg.k));
  // scratch layout: counter | job record | result record | read | matrices | graph block
  const uint64_t off_seq = 768;
  const uint64_t off_mat = RoundUp(off_seq + stride + 16, kAlign);
  const uint64_t mat_bytes = JobBytes(E->prm, V, static_cast<uint32_t>(stride), sh.bits, cfg);
  const uint64_t off_graph = RoundUp(off_mat + mat_bytes, kAlign);
  if (off_graph + (9ull * V + 4) * 4 > kScratchBytes) {
    return Fail(SPOA_HIP_NO_MEMORY,
                "[spoa::hip] the alignment needs " + std::to_string(off_graph + (9ull * V + 4) * 4) +
                    " bytes of device scratch, more than the " + std::to_string(kScratchBytes >> 20) +
                    " MiB the no-allocation entry point has");
  }
  // (its one job is never split: a workgroup runs the row tiles in turn)
  KernelFn fn = PickKernel(E->prm.subtype, sh.bits, sh.lanes, cfg.k, cfg.waves, false, E->sisd);
  if (!fn) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] unsupported launch configuration");
  spoa_hip_dev::NoAllocPackArgs pa{};
  pa.row_char = graph->row_char;
  pa.node_id = graph->node_id;
  pa.pred_off = graph->pred_off;
  pa.pred_rank = graph->pred_rank;
  pa.is_sink = graph->is_sink;
  pa.seq = seq;
  pa.V = V;
  pa.n = len;
  pa.W = W;
  pa.stride = static_cast<uint32_t>(stride);
  pa.cap = cap;
  pa.type = E->prm.type;
  pa.subtype = E->prm.subtype;
  pa.g = E->prm.g;
  pa.e = E->prm.e;
  pa.q = E->prm.q;
  pa.c = E->prm.c;
  pa.counter = reinterpret_cast<uint32_t*>(base);
  pa.job = reinterpret_cast<DevJob*>(base + 256);
  pa.res = reinterpret_cast<int32_t*>(base + 512);
  pa.seq_out = base + off_seq;
  pa.mat_addr = reinterpret_cast<uint64_t>(base + off_mat);
  pa.graph = reinterpret_cast<uint32_t*>(base + off_graph);
  pa.graph_cap = (kScratchBytes - off_graph) / 4;
  spoa_hip_dev::NoAllocPack<<<1, 256, 0, st>>>(pa);
  if (!HipOk(hipGetLastError(), "kernel launch")) return SPOA_HIP_DEVICE_ERROR;
  KArgs args{};
  args.prm = MakeParams(E->prm, sh.bits, sh.lanes, cfg.k, E->sisd);
  args.jobs = pa.job;
  args.n_jobs = 1;
  args.queue = pa.counter;
  args.graphs = pa.graph;
  args.seqs = pa.seq_out;
  args.pairs = pairs;
  args.results = pa.res;
  const int threads = 64 * cfg.waves;
  args.lds_bytes = 8u * static_cast<uint32_t>(cfg.k) * static_cast<uint32_t>(threads);
  args.prio_jobs = 0;
  args.phases = E->phases;
  fn<<<1, threads, args.lds_bytes, st>>>(args);
  if (!HipOk(hipGetLastError(), "kernel launch")) return SPOA_HIP_DEVICE_ERROR;
  spoa_hip_dev::NoAllocFinish<<<1, 256, 0, st>>>(pa.res, pairs, cap, sh.width, result);
  return HipOk(hipGetLastError(), "kernel launch") ? SPOA_HIP_OK : SPOA_HIP_DEVICE_ERROR;
}

int spoa_hip_align_batch(spoa_hip_engine* E, spoa_hip_job* jobs, uint32_t n_jobs) {
  if (!E || (!jobs && n_jobs)) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null argument");
  if (!HipOk(hipSetDevice(E->device), "hipSetDevice")) return SPOA_HIP_DEVICE_ERROR;
  return RunBatch(E, jobs, n_jobs, true, false);
}

int spoa_hip_align_batch_submit(spoa_hip_engine* E, spoa_hip_job* jobs, uint32_t n_jobs) {
  if (!E || (!jobs && n_jobs)) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null argument");
  if (!HipOk(hipSetDevice(E->device), "hipSetDevice")) return SPOA_HIP_DEVICE_ERROR;
  const int rc = RunBatch(E, jobs, n_jobs, true, true);
  return rc == SPOA_HIP_DEVICE_ERROR || rc == SPOA_HIP_NO_MEMORY ? rc : SPOA_HIP_OK;
}

int spoa_hip_align_batch_wait(spoa_hip_engine* E) {
  if (!E) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null engine");
  if (!HipOk(hipSetDevice(E->device), "hipSetDevice")) return SPOA_HIP_DEVICE_ERROR;
  const std::string before = g_last_error;
  const int rc = RunPassFinish(E);
  if (rc != SPOA_HIP_OK) return rc;
  if (E->pending.first != SPOA_HIP_OK) g_last_error = before;
  return E->pending.first;
}

int spoa_hip_align_batch_ready(spoa_hip_engine* E) {
  if (!E) return -Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null engine");
  if (!E->pending.active) return 1;
  const hipError_t q = E->graph_pass ? hipEventQuery(E->ev_kdone) : hipStreamQuery(E->copy);
  if (q == hipErrorNotReady) return 0;
  (void)hipGetLastError();
  return 1;  // finished, or failed: wait reports which
}

int spoa_hip_last_stats(spoa_hip_engine* E, spoa_hip_stats* out) {
  if (!E || !out) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null argument");
  *out = E->stats;
  return SPOA_HIP_OK;
}

int spoa_hip_config_split(spoa_hip_engine* E, uint32_t len, int32_t width_bits, int32_t batch, int32_t* k,
                          int32_t* waves, int32_t* split) {
  if (!E || !k || !waves || !split || (width_bits != 16 && width_bits != 32)) {
    return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] invalid argument");
  }
  const Cfg c = ChooseCfg(E, len, width_bits, batch != 0);
  *k = c.k;
  *waves = c.waves;
  *split = c.split;
  return SPOA_HIP_OK;
}

int spoa_hip_config(spoa_hip_engine* E, uint32_t len, int32_t width_bits, int32_t batch, int32_t* k, int32_t* waves) {
  int32_t split = 0;
  return spoa_hip_config_split(E, len, width_bits, batch, k, waves, &split);
}

int spoa_hip_set_config_split(spoa_hip_engine* E, int32_t width_bits, int32_t batch, uint32_t max_width, int32_t k,
                              int32_t waves, int32_t split) {
  if (!E || (width_bits != 16 && width_bits != 32)) {
    return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] invalid argument");
  }
  auto& table = E->overrides[{width_bits, batch ? 1 : 0}];
  if (k == 0) {
    table.clear();
    return SPOA_HIP_OK;
  }
  if ((k != 8 && k != 16 && k != 32) || waves < 1 || waves > kMaxGroupWaves ||
      (E->prm.subtype == 2 && width_bits == 16 && k == 8) || (k == 32 && waves != 1)) {
    return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] unsupported launch configuration");
  }
  Cfg c;
  c.k = k;
  c.waves = waves;
  c.split = split != 0 ? 1 : 0;
  table[max_width] = c;
  return SPOA_HIP_OK;
}

int spoa_hip_set_config(spoa_hip_engine* E, int32_t width_bits, int32_t batch, uint32_t max_width, int32_t k,
                        int32_t waves) {
  return spoa_hip_set_config_split(E, width_bits, batch, max_width, k, waves, 0);
}

int spoa_hip_set_wait_limit(spoa_hip_engine* E, uint32_t ticks) {
  if (!E) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null engine");
  E->wait_ticks = ticks;
  return SPOA_HIP_OK;
}

int spoa_hip_share_compute(spoa_hip_engine* E, spoa_hip_engine* owner) {
  if (!E) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null engine");
  if (owner && owner->device != E->device) {
    return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] engines on different devices cannot share a stream");
  }
  if (E->pending.active) (void)RunPassFinish(E);
  if (owner && owner->compute == owner->stream) {
    // the owner's kernels move off its copy stream, so no engine's copies
    // wait behind another engine's kernels
    if (owner->pending.active) (void)RunPassFinish(owner);
    if (!HipOk(hipSetDevice(owner->device), "hipSetDevice") ||
        !HipOk(hipStreamCreateWithFlags(&owner->compute_own, hipStreamNonBlocking), "hipStreamCreate")) {
      return SPOA_HIP_DEVICE_ERROR;
    }
    owner->compute = owner->compute_own;
  }
  E->compute = owner ? owner->compute : E->stream;
  // side launches on the engine's own stream (idle but for its copies), so
  // the engines sharing the compute stream each have their own
  if (owner) owner->side = owner->compute != owner->stream ? owner->stream : nullptr;
  E->side = E->compute != E->stream ? E->stream : nullptr;
  return SPOA_HIP_OK;
}

int spoa_hip_share_stream(spoa_hip_engine* E, spoa_hip_engine* owner) {
  if (!E) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null engine");
  if (owner && owner->device != E->device) {
    return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] engines on different devices cannot share a stream");
  }
  if (E->pending.active) (void)RunPassFinish(E);
  E->compute = E->copy = owner ? owner->stream : E->stream;
  // everything of this engine on the owner's one stream
  E->side = owner ? owner->stream : nullptr;
  return SPOA_HIP_OK;
}

int spoa_hip_set_graph_mode(spoa_hip_engine* E, int32_t on) {
  if (!E) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null engine");
  if (E->pending.active) (void)RunPassFinish(E);
  E->graph_mode = on != 0;
  return SPOA_HIP_OK;
}

int spoa_hip_set_memory_budget(spoa_hip_engine* E, uint64_t bytes) {
  if (!E) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null engine");
  E->budget = bytes;
  return SPOA_HIP_OK;
}

uint64_t spoa_hip_memory_budget(spoa_hip_engine* E) { return E ? E->budget : 0; }

uint64_t spoa_hip_memory_in_use(spoa_hip_engine* E) { return E ? E->arena_bytes : 0; }

uint64_t spoa_hip_job_footprint(spoa_hip_engine* E, uint32_t V, uint32_t len, int32_t batch) {
  if (!E || V == 0 || len == 0) return 0;
  const Shape sh = JobShape(E->prm, len, V, E->sisd);
  if (sh.width == 0) return 0;
  const Cfg cfg = ChooseCfg(E, len, sh.cfgw, batch != 0);
  const int L = sh.lanes;
  const uint32_t W = (len + L - 1) / L * L;
  const uint32_t stride = static_cast<uint32_t>(RoundUp(W, static_cast<uint64_t>(cfg.k)));
  return JobBytes(E->prm, V, stride, sh.bits, cfg);
}

}  // extern "C"
