// Device-resident POA windows (spoa_hip_windows_* in include/spoa_hip.h).
//
// Every window's graph lives in device memory for the whole run.  A round
// aligns the next read of every live window with the engine's batched
// alignment kernels, whose graph blocks the previous round's merge wrote,
// then one wavefront per window merges the alignment into the window's
// graph (AddAlignment), sorts it (TopologicalSort) and writes the next
// round's graph block (window_graph.hpp).  The host only plans the round's
// launches from a 32-byte summary per window (nodes, edges, status) that
// comes back at the end of every round.  Consensus and MSA are generated on
// the device at the end.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "engine.hpp"
#include "pass_split.hpp"
#include "spoa_hip.h"
#include "window_graph.hpp"

using namespace spoa_hip_host;
namespace wg = spoa_amd::wg;

namespace spoa_hip_dev {

// One wavefront as the team of window_graph.hpp.
struct WaveTeam {
  uint32_t lane;
  static constexpr uint32_t size = 64;
  __device__ void sync() const { __syncthreads(); }
  __device__ uint32_t scan(uint32_t v, uint32_t* total) const {
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (lane >= static_cast<uint32_t>(d)) x += y;
    }
    *total = __shfl(x, 63, 64);
    return x - v;
  }
  __device__ bool any(bool b) const { return __ballot(b) != 0; }
  // lane 0's value, held in a scalar register: branches on it are scalar
  __device__ uint32_t uniform(uint32_t v) const {
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
  }
  // the lanes where b holds, as a bit mask
  __device__ uint64_t ballot(bool b) const { return __ballot(b); }
  // set bits of m below this lane
  __device__ uint32_t below(uint64_t m) const {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
  }
  // lane l's value (l uniform)
  __device__ uint32_t readlane(uint32_t v, uint32_t l) const {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(l)));
  }
  // the minimum over lanes 0-15, uniform: a DPP prefix minimum within the
  // first row of 16 lanes, read from its last lane
  __device__ uint32_t min16(uint32_t v) const {
    int32_t x = static_cast<int32_t>(v);
    x = static_cast<int32_t>(min(static_cast<uint32_t>(x), static_cast<uint32_t>(__builtin_amdgcn_update_dpp(x, x, 0x111, 0xf, 0xf, false))));
    x = static_cast<int32_t>(min(static_cast<uint32_t>(x), static_cast<uint32_t>(__builtin_amdgcn_update_dpp(x, x, 0x112, 0xf, 0xf, false))));
    x = static_cast<int32_t>(min(static_cast<uint32_t>(x), static_cast<uint32_t>(__builtin_amdgcn_update_dpp(x, x, 0x114, 0xf, 0xf, false))));
    x = static_cast<int32_t>(min(static_cast<uint32_t>(x), static_cast<uint32_t>(__builtin_amdgcn_update_dpp(x, x, 0x118, 0xf, 0xf, false))));
    return readlane(static_cast<uint32_t>(x), 15);
  }
  __device__ uint32_t min_all(uint32_t v) const {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = min(v, static_cast<uint32_t>(__shfl_xor(v, d, 64)));
    return v;
  }
  __device__ uint32_t max_all(uint32_t v) const {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, static_cast<uint32_t>(__shfl_xor(v, d, 64)));
    return v;
  }
  __device__ uint32_t sum_all(uint32_t v) const {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += static_cast<uint32_t>(__shfl_xor(v, d, 64));
    return v;
  }
};

// Where a window lives.
struct WinInfo {
  uint64_t region;      // byte offset of its region in the window buffer
  wg::WinCaps caps;
  uint32_t block_off;   // 32-bit words of its graph block into the block buffer
  uint32_t pad[3];
};

// What the host reads back after every round (and after the finish).
struct WinSummary {
  uint32_t V, E, A, R;
  uint32_t nslots, slot_chars;
  int32_t status;
  uint32_t aux;  // finish: consensus length (MSA width in aux2)
};

// One window's merge of a round.
struct MergeItem {
  uint32_t window;
  int32_t job_f, job_r;  // jobs of the read's two strands in the round's table (-1: none)
  uint32_t read;         // read index in the window
  uint32_t len;
  uint32_t seq_f, seq_r;  // byte offsets of the read and its reverse complement
  uint32_t qual;          // byte offset of its qualities (when the run has them)
};

struct MergeArgs {
  const MergeItem* items;
  const WinInfo* info;
  uint8_t* windows;
  uint32_t* blocks;
  const DevJob* jobs;
  const int32_t* results;  // 4 per job: n_pairs, score, has_score, status
  const int32_t* pairs;
  const uint8_t* seqs;
  const uint8_t* quals;    // null: weight 1
  WinSummary* summary;
  int32_t type, subtype, g, e, q, c;
  uint32_t lds_bytes;
};

__device__ void Summarise(const wg::WinView& g, WinSummary* out, uint32_t A) {
  out->V = g.hdr->V;
  out->E = g.hdr->E;
  out->A = A;
  out->R = g.hdr->R;
  out->nslots = g.hdr->nslots;
  out->slot_chars = g.hdr->slot_chars;
  out->status = g.hdr->status;
}

// Round r: one wave a window -- the strand with the better score
// (src/main.cpp:285-302, forward on ties), AddAlignment, TopologicalSort,
// the next graph block.
__global__ void __launch_bounds__(64) MergeKernel(MergeArgs a) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
  const WaveTeam t{threadIdx.x};
  const MergeItem it = a.items[blockIdx.x];
  const WinInfo wi = a.info[it.window];
  const wg::WinView g = wg::MakeWinView(a.windows + wi.region, wi.caps);
  WinSummary* sum = a.summary + it.window;
  if (g.hdr->status != wg::kWinOk) return;
  // the alignment: no job (empty graph or read) is the empty alignment
  const int32_t* pairs = nullptr;
  uint32_t np = 0;
  bool reverse = false;
  int32_t failed = 0;
  if (it.job_f >= 0) {
    const int32_t* rf = a.results + 4 * it.job_f;
    int32_t sf = rf[2] ? rf[1] : 0;
    failed |= rf[3];
    const int32_t* use = rf;
    int32_t job = it.job_f;
    if (it.job_r >= 0) {
      const int32_t* rr = a.results + 4 * it.job_r;
      const int32_t sr = rr[2] ? rr[1] : 0;
      failed |= rr[3];
      if (!(sf >= sr)) {
        use = rr;
        job = it.job_r;
        reverse = true;
      }
    }
    np = static_cast<uint32_t>(use[0]);
    pairs = a.pairs + 2ull * a.jobs[job].pair_off;
  } else if (it.job_r >= 0) {
    failed = 1;  // a reverse strand job without the forward one never happens
  }
  if (failed) {
    if (t.lane == 0) {
      g.hdr->status = wg::kWinAlignFailed;
      sum->status = wg::kWinAlignFailed;
    }
    return;
  }
  if (t.lane == 0) g.reversed[it.read] = reverse ? 1 : 0;
  wg::AlignIn in;
  in.pairs = pairs;
  in.np = np;
  in.seq = a.seqs + (reverse ? it.seq_r : it.seq_f);
  in.len = it.len;
  in.qual = a.quals ? a.quals + it.qual : nullptr;
  in.qual_reversed = reverse;
  const int32_t st = wg::AddAlignment(t, g, in);
  uint32_t A = 0;
  for (uint32_t v = t.lane; v < g.hdr->V; v += t.size) A += g.al_cnt[v];
  A = t.sum_all(A);
  if (st != wg::kWinOk) {
    if (t.lane == 0) g.hdr->status = st;
    t.sync();
    if (t.lane == 0) Summarise(g, sum, A);
    return;
  }
  if (it.len == 0) {
    if (t.lane == 0) Summarise(g, sum, A);
    return;
  }
  const uint32_t V = g.hdr->V, E = g.hdr->E;
  const bool nw = a.type == 1;
  bool sorted = false;
  uint32_t* const block = a.blocks + wi.block_off;
  if (V < 65535 && E < 65535 && A < 65535) {
    // 16-bit lists in LDS, the stack taking what is left
    const uint64_t fixed = wg::SortSpaceBytes(V, E, A, 0, nw, 2);
    if (fixed + 2 * 64 + 16 <= a.lds_bytes) {
      const uint64_t room = (a.lds_bytes - fixed - 16) / 2, full = static_cast<uint64_t>(V) + E + A + 1;
      const uint32_t K = static_cast<uint32_t>(room < full ? room : full);
      const auto sp = wg::CarveSortSpace<uint16_t>(lds, V, E, A, K, nw);
      sorted = wg::TopologicalSort(t, g, sp);
      if (sorted) wg::BuildBlock(t, g, sp, block, a.type, a.subtype, a.g, a.e, a.q, a.c);
    }
  }
  if (!sorted) {
    const uint64_t cap_words = wg::MakeLayout(wi.caps).scratch_words;
    const uint64_t need = (wg::SortSpaceBytes(V, E, A, static_cast<uint64_t>(V) + E + A + 1, nw, 4) + 3) / 4;
    if (need > cap_words) {
      if (t.lane == 0) {
        g.hdr->status = wg::kWinCapacity;
        Summarise(g, sum, A);
      }
      return;
    }
    const auto sp = wg::CarveSortSpace<uint32_t>(g.scratch, V, E, A, V + E + A + 1, nw);
    wg::TopologicalSort(t, g, sp);
    wg::BuildBlock(t, g, sp, block, a.type, a.subtype, a.g, a.e, a.q, a.c);
  }
  if (t.lane == 0) Summarise(g, sum, A);
}

__global__ void __launch_bounds__(64) InitWindowsKernel(const WinInfo* info, uint8_t* windows, WinSummary* summary,
                                                        uint32_t n) {
  const WaveTeam t{threadIdx.x};
  for (uint32_t w = blockIdx.x; w < n; w += gridDim.x) {
    const wg::WinView g = wg::MakeWinView(windows + info[w].region, info[w].caps);
    wg::InitWindow(t, g);
    if (t.lane == 0) summary[w] = WinSummary{};
    t.sync();
  }
}

// The end of a run, first half: every window's consensus node ids (into its
// graph block, which no alignment reads any more) and MSA width.
__global__ void __launch_bounds__(64) ConsensusKernel(const WinInfo* info, uint8_t* windows, uint32_t* blocks,
                                                      WinSummary* summary, uint32_t n) {
  const WaveTeam t{threadIdx.x};
  const uint32_t w = blockIdx.x;
  if (w >= n) return;
  const wg::WinView g = wg::MakeWinView(windows + info[w].region, info[w].caps);
  if (g.hdr->status != wg::kWinOk) return;
  uint32_t* cons = blocks + info[w].block_off;
  const uint32_t len = wg::Consensus(t, g, g.scratch, cons);
  const uint32_t width = wg::MsaColumns(t, g, g.scratch);
  if (t.lane == 0) {
    summary[w].aux = len;
    summary[w].A = width;
    summary[w].V = g.hdr->V;
    summary[w].E = g.hdr->nseq;
    summary[w].status = g.hdr->status;
  }
}

// Second half: consensus bases and MSA rows (sequences in insertion order)
// packed at the host's offsets.
__global__ void __launch_bounds__(64) OutputKernel(const WinInfo* info, uint8_t* windows, const uint32_t* blocks,
                                                   const uint64_t* cons_off, const uint64_t* msa_off, int32_t msa,
                                                   uint8_t* out, uint32_t n) {
  const WaveTeam t{threadIdx.x};
  const uint32_t w = blockIdx.x;
  if (w >= n) return;
  const wg::WinView g = wg::MakeWinView(windows + info[w].region, info[w].caps);
  if (g.hdr->status != wg::kWinOk) return;
  const uint32_t* cons = blocks + info[w].block_off;
  const uint32_t clen = static_cast<uint32_t>(cons_off[w + 1] - cons_off[w]);
  for (uint32_t k = t.lane; k < clen; k += t.size) out[cons_off[w] + k] = g.hdr->decoder[g.code[cons[k]]];
  if (!msa) return;
  const uint32_t width = wg::MsaColumns(t, g, g.scratch);
  wg::WriteMsa(t, g, g.scratch, width, cons, 0, out + msa_off[w]);
}

}  // namespace spoa_hip_dev

using spoa_hip_dev::MergeItem;
using spoa_hip_dev::WinInfo;
using spoa_hip_dev::WinSummary;

struct spoa_hip_windows {
  spoa_hip_engine* E = nullptr;
  uint32_t n = 0;
  bool strand = false, quals = false;
  uint32_t rounds = 0;
  std::vector<uint32_t> nreads;            // per window
  std::vector<uint32_t> read_len;          // per read (window-major)
  std::vector<uint32_t> read_first;        // index of each window's first read
  std::vector<uint32_t> seq_f, seq_r, qual;  // device byte offsets per read
  std::vector<WinInfo> info;
  DevBuf d_windows, d_blocks, d_seqs, d_quals, d_info, d_summary;
  DevBuf d_table, d_res, d_pairs, d_counter, d_out, d_offs;
  PinBuf h_table, h_summary, h_out, h_offs, h_stage;
  std::vector<WinSummary> state;  // per window, as of the last finished round
  // the submitted round
  bool pending = false;
  uint32_t round = 0;
  std::vector<uint32_t> live;  // windows merged in the round
  std::vector<hipEvent_t> ev_s, ev_e;  // DP launches
  hipEvent_t ev_m0 = nullptr, ev_m1 = nullptr, ev_done = nullptr;
  size_t launches = 0;
  std::vector<int> launch_width;
  spoa_hip_stats stats{};
  uint64_t round_dp_bytes16 = 0, round_dp_bytes32 = 0;
  // finish results
  std::vector<uint64_t> cons_off, msa_off;
  std::vector<uint32_t> width, nseq;
  bool finished = false, msa_done = false;
  std::string last_error;
};

namespace {

// The window set's buffers and events, in the order of the engine's
// WinBuffers cache.
std::array<DevBuf*, 12> DevBufsOf(spoa_hip_windows* X) {
  return {&X->d_windows, &X->d_blocks, &X->d_seqs, &X->d_quals, &X->d_info, &X->d_summary,
          &X->d_table,   &X->d_res,    &X->d_pairs, &X->d_counter, &X->d_out, &X->d_offs};
}
std::array<PinBuf*, 5> PinBufsOf(spoa_hip_windows* X) {
  return {&X->h_table, &X->h_summary, &X->h_out, &X->h_offs, &X->h_stage};
}

// Takes over the buffers the engine kept from its last window set.
void AdoptBuffers(spoa_hip_windows* X) {
  auto& c = X->E->win_bufs;
  if (!c.held) return;
  auto d = DevBufsOf(X);
  for (size_t i = 0; i < d.size(); ++i) {
    *d[i] = c.d[i];
    c.d[i] = DevBuf{};
  }
  auto h = PinBufsOf(X);
  for (size_t i = 0; i < h.size(); ++i) {
    *h[i] = c.h[i];
    c.h[i] = PinBuf{};
  }
  X->ev_s.swap(c.ev_s);
  X->ev_e.swap(c.ev_e);
  c.ev_s.clear();
  c.ev_e.clear();
  X->ev_m0 = c.ev_m0;
  X->ev_m1 = c.ev_m1;
  X->ev_done = c.ev_done;
  c.ev_m0 = c.ev_m1 = c.ev_done = nullptr;
  c.held = false;
}

int WinFail(spoa_hip_windows* X, int code, const std::string& msg) {
  X->last_error = msg;
  return Fail(code, msg);
}

const char* WinStatusText(int32_t st) {
  switch (st) {
    case wg::kWinAlignFailed: return "[spoa::hip] the alignment of a window failed";
    case wg::kWinInvalidAlignment: return "[spoa::Graph::AddAlignment] error: invalid alignment";
    case wg::kWinMissingSequence: return "[spoa::Graph::AddAlignment] error: missing sequence in alignment";
    case wg::kWinGroupTooLarge: return "[spoa::hip] an aligned group of more than 8 codes";
    case wg::kWinGroupRevisit: return "[spoa::hip] an alignment through two nodes of one aligned group";
    case wg::kWinCapacity: return "[spoa::hip] a window graph outgrew its device region";
    default: return "[spoa::hip] window failed";
  }
}

// The reverse complement of the strand-ambiguous mode: upper-case, then the
// Watson-Crick / IUPAC complement (the rule of spoa_reverse_complement).
char Complement(char b) {
  switch (b) {
    case 'A': case 'a': return 'T';
    case 'C': case 'c': return 'G';
    case 'G': case 'g': return 'C';
    case 'T': case 't': case 'U': case 'u': return 'A';
    case 'R': case 'r': return 'Y';
    case 'Y': case 'y': return 'R';
    case 'K': case 'k': return 'M';
    case 'M': case 'm': return 'K';
    case 'B': case 'b': return 'V';
    case 'V': case 'v': return 'B';
    case 'D': case 'd': return 'H';
    case 'H': case 'h': return 'D';
    case 'S': case 's': return 'S';
    case 'W': case 'w': return 'W';
    default: return b;
  }
}

// LDS of the merge launch: the largest 16-bit sort space among its windows
// (their sizes after the merge bounded by the read) with a stack of
// kMergeStack entries, at most kMergeLdsMax; larger windows, and a sort
// whose stack outgrows what is left, sort in device memory.  (The deepest
// stacks measured: 327 entries on overlap windows of 500 bp, whose reads'
// overhangs become chains of new nodes; about 10 on the others.)  The LDS
// a workgroup takes sets how many windows sort side by side on a CU.
constexpr uint64_t kMergeLdsMax = 160 * 1024;
constexpr uint64_t kMergeStack = 512;
uint32_t MergeLds(const spoa_hip_windows* X, uint32_t r, bool nw) {
  uint64_t best = 4096;
  for (uint32_t w : X->live) {
    const WinSummary& s = X->state[w];
    const uint64_t len = X->read_len[X->read_first[w] + r];
    const uint64_t V = s.V + len, E = s.E + len, A = s.A + 2 * len;
    if (V >= 65535 || E >= 65535 || A >= 65535) continue;
    const uint64_t b = wg::SortSpaceBytes(V, E, A, kMergeStack, nw, 2) + 64;
    best = std::max(best, std::min<uint64_t>(b, kMergeLdsMax));
  }
  return static_cast<uint32_t>((best + 1023) / 1024 * 1024);
}

}  // namespace

extern "C" {

int spoa_hip_windows_create(spoa_hip_engine* E, uint32_t n_windows, const uint32_t* win_read_off,
                            const uint64_t* read_off, const char* seqs, const char* quals, int32_t strand,
                            spoa_hip_windows** out) {
  if (!E || !out || (n_windows && (!win_read_off || !read_off || !seqs))) {
    return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null argument");
  }
  if (!HipOk(hipSetDevice(E->device), "hipSetDevice")) return SPOA_HIP_DEVICE_ERROR;
  if (E->pending.active) {
    return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] the engine has a batch in flight");
  }
  auto* X = new spoa_hip_windows();
  X->E = E;
  AdoptBuffers(X);
  X->n = n_windows;
  X->strand = strand != 0;
  X->quals = quals != nullptr;
  X->nreads.resize(n_windows);
  X->read_first.resize(n_windows + 1);
  X->info.resize(n_windows);
  X->state.assign(n_windows, WinSummary{});
  const uint32_t n_reads = n_windows ? win_read_off[n_windows] - win_read_off[0] : 0;
  X->read_len.resize(n_reads);
  X->seq_f.resize(n_reads);
  X->seq_r.resize(n_reads);
  X->qual.resize(n_reads);
  // layout: windows' regions, graph blocks, reads (forward, then reverse
  // complements), each read padded for the alignment kernels' row reads
  uint64_t region = 0, blocks = 0, sbytes = 0;
  for (uint32_t w = 0; w < n_windows; ++w) {
    const uint32_t a = win_read_off[w], b = win_read_off[w + 1];
    X->nreads[w] = b - a;
    X->read_first[w] = a - win_read_off[0];
    X->rounds = std::max(X->rounds, b - a);
    uint64_t total = 0;
    for (uint32_t k = a; k < b; ++k) {
      const uint64_t len = read_off[k + 1] - read_off[k];
      if (len > static_cast<uint64_t>(std::numeric_limits<int32_t>::max())) {
        delete X;
        return Fail(SPOA_HIP_TOO_LARGE, "[spoa::SimdAlignmentEngine::Align] error: too large sequence!");
      }
      X->read_len[k - win_read_off[0]] = static_cast<uint32_t>(len);
      total += len;
    }
    wg::WinCaps caps{};
    caps.V = static_cast<uint32_t>(std::max<uint64_t>(total, 1));
    caps.E = caps.V;
    caps.L = caps.V;
    caps.S = std::max<uint32_t>(b - a, 1);
    X->info[w].region = region;
    X->info[w].caps = caps;
    X->info[w].block_off = static_cast<uint32_t>(blocks);
    region += wg::MakeLayout(caps).total;
    blocks += wg::BlockWords(caps.V, caps.E);
    if (blocks > 0xffffffffull) {
      delete X;
      return Fail(SPOA_HIP_NO_MEMORY, "[spoa::hip] the windows' graph blocks exceed 16 GiB; run fewer windows");
    }
  }
  X->read_first[n_windows] = n_reads;
  const int copies = X->strand ? 2 : 1;
  for (int d = 0; d < copies; ++d) {
    for (uint32_t k = 0; k < n_reads; ++k) {
      (d == 0 ? X->seq_f : X->seq_r)[k] = static_cast<uint32_t>(sbytes);
      if (d == 0) X->qual[k] = static_cast<uint32_t>(sbytes);
      sbytes += (static_cast<uint64_t>(X->read_len[k]) + 64 + 15) / 16 * 16;
      if (sbytes > 0xffffffffull) {
        delete X;
        return Fail(SPOA_HIP_NO_MEMORY, "[spoa::hip] the windows' reads exceed 4 GiB; run fewer windows");
      }
    }
  }
  if (!X->strand) X->seq_r = X->seq_f;
  const uint64_t qbytes = X->quals ? sbytes : 0;
  if (!X->d_windows.Reserve(std::max<uint64_t>(region, 256), "[spoa::hip] insufficient memory for the windows") ||
      !X->d_blocks.Reserve(std::max<uint64_t>(blocks * 4, 256), "[spoa::hip] insufficient memory for the windows") ||
      !X->d_seqs.Reserve(std::max<uint64_t>(sbytes, 256), "[spoa::hip] insufficient memory for the windows") ||
      (qbytes && !X->d_quals.Reserve(qbytes, "[spoa::hip] insufficient memory for the windows")) ||
      !X->d_info.Reserve(std::max<size_t>(n_windows, 1) * sizeof(WinInfo), "[spoa::hip] insufficient memory") ||
      !X->d_summary.Reserve(std::max<size_t>(n_windows, 1) * sizeof(WinSummary), "[spoa::hip] insufficient memory") ||
      !X->d_counter.Reserve(256 * sizeof(uint32_t), "[spoa::hip] insufficient memory") ||
      !X->h_summary.Reserve(std::max<size_t>(n_windows, 1) * sizeof(WinSummary))) {
    const std::string msg = spoa_hip_last_error();
    spoa_hip_windows_destroy(X);
    return Fail(SPOA_HIP_NO_MEMORY, msg);
  }
  // reads (and reverse complements, qualities) in one staging copy
  {
    PinBuf& stage = X->h_stage;
    if (!stage.Reserve(std::max<uint64_t>(sbytes, 256))) {
      spoa_hip_windows_destroy(X);
      return SPOA_HIP_DEVICE_ERROR;
    }
    uint8_t* h = stage.At<uint8_t>(0);
    std::memset(h, 0, sbytes);
    for (uint32_t k = 0; k < n_reads; ++k) {
      const char* s = seqs + read_off[win_read_off[0] + k];
      std::memcpy(h + X->seq_f[k], s, X->read_len[k]);
      if (X->strand) {
        const uint32_t len = X->read_len[k];
        for (uint32_t i = 0; i < len; ++i) h[X->seq_r[k] + len - 1 - i] = static_cast<uint8_t>(Complement(s[i]));
      }
    }
    bool ok = HipOk(hipMemcpyAsync(X->d_seqs.p, h, sbytes, hipMemcpyHostToDevice, E->stream), "copy in") &&
              HipOk(hipStreamSynchronize(E->stream), "copy in");
    if (ok && qbytes) {
      std::memset(h, 0, sbytes);
      for (uint32_t k = 0; k < n_reads; ++k) {
        std::memcpy(h + X->qual[k], quals + read_off[win_read_off[0] + k], X->read_len[k]);
      }
      ok = HipOk(hipMemcpyAsync(X->d_quals.p, h, qbytes, hipMemcpyHostToDevice, E->stream), "copy in") &&
           HipOk(hipStreamSynchronize(E->stream), "copy in");
    }
    if (ok) {
      ok = HipOk(hipMemcpyAsync(X->d_info.p, X->info.data(), n_windows * sizeof(WinInfo), hipMemcpyHostToDevice,
                                E->stream),
                 "copy in");
    }
    if (ok && n_windows) {
      spoa_hip_dev::InitWindowsKernel<<<std::min<uint32_t>(n_windows, 65535), 64, 0, E->stream>>>(
          static_cast<const WinInfo*>(X->d_info.p), static_cast<uint8_t*>(X->d_windows.p),
          static_cast<WinSummary*>(X->d_summary.p), n_windows);
      ok = HipOk(hipGetLastError(), "kernel launch");
    }
    if (ok) ok = HipOk(hipStreamSynchronize(E->stream), "window init");
    if (!ok) {
      spoa_hip_windows_destroy(X);
      return SPOA_HIP_DEVICE_ERROR;
    }
  }
  if ((!X->ev_m0 && !HipOk(hipEventCreate(&X->ev_m0), "hipEventCreate")) ||
      (!X->ev_m1 && !HipOk(hipEventCreate(&X->ev_m1), "hipEventCreate")) ||
      (!X->ev_done && !HipOk(hipEventCreateWithFlags(&X->ev_done, hipEventDisableTiming), "hipEventCreate"))) {
    spoa_hip_windows_destroy(X);
    return SPOA_HIP_DEVICE_ERROR;
  }
  // the DP arena shares what is left of the device with the windows
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
    const uint64_t cap = static_cast<uint64_t>(static_cast<double>(free_b) * 0.8) + E->arena_bytes;
    if (E->budget > cap) E->budget = cap;
  }
  *out = X;
  return SPOA_HIP_OK;
}

int spoa_hip_windows_destroy(spoa_hip_windows* X) {
  if (!X) return SPOA_HIP_OK;
  spoa_hip_engine* E = X->E;
  (void)hipSetDevice(E->device);
  // nothing of the set may be in flight when the next set takes its buffers
  (void)hipStreamSynchronize(E->stream);
  // the engine keeps the buffers for its next window set (one set's worth:
  // what it held before is freed)
  ReleaseWinBuffers(E);
  auto& c = E->win_bufs;
  auto d = DevBufsOf(X);
  for (size_t i = 0; i < d.size(); ++i) c.d[i] = *d[i];
  auto h = PinBufsOf(X);
  for (size_t i = 0; i < h.size(); ++i) c.h[i] = *h[i];
  c.ev_s.swap(X->ev_s);
  c.ev_e.swap(X->ev_e);
  c.ev_m0 = X->ev_m0;
  c.ev_m1 = X->ev_m1;
  c.ev_done = X->ev_done;
  c.held = true;
  delete X;
  return SPOA_HIP_OK;
}

uint32_t spoa_hip_windows_rounds(spoa_hip_windows* X) { return X ? X->rounds : 0; }

int spoa_hip_windows_submit(spoa_hip_windows* X, uint32_t r) {
  if (!X) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null argument");
  if (X->pending) return WinFail(X, SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] a round is in flight");
  if (X->finished) return WinFail(X, SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] the windows are finished");
  spoa_hip_engine* E = X->E;
  if (!HipOk(hipSetDevice(E->device), "hipSetDevice")) return SPOA_HIP_DEVICE_ERROR;
  const spoa_hip_params& P = E->prm;
  const auto t0 = std::chrono::steady_clock::now();
  X->round = r;
  X->live.clear();
  for (uint32_t w = 0; w < X->n; ++w) {
    if (r < X->nreads[w] && X->state[w].status == wg::kWinOk) X->live.push_back(w);
  }
  // jobs: read r of every live window with a graph, both strands in strand mode
  struct JobRef {
    uint32_t window, strand;
  };
  std::vector<Plan> plans;
  std::vector<JobRef> refs;
  uint32_t cfg_len[2] = {0, 0};
  for (uint32_t w : X->live) {
    const uint32_t len = X->read_len[X->read_first[w] + r];
    const uint32_t V = X->state[w].V;
    if (V == 0 || len == 0) continue;
    const Shape sh = JobShape(P, len, V, E->sisd);
    if (sh.width == 0) return WinFail(X, SPOA_HIP_OVERFLOW, E->msg->overflow);
    cfg_len[sh.cfgw == 32] = std::max(cfg_len[sh.cfgw == 32], len);
  }
  for (uint32_t w : X->live) {
    const uint32_t len = X->read_len[X->read_first[w] + r];
    const uint32_t V = X->state[w].V;
    if (V == 0 || len == 0) continue;
    const Shape sh = JobShape(P, len, V, E->sisd);
    for (uint32_t d = 0; d < (X->strand ? 2u : 1u); ++d) {
      Plan pl{};
      pl.job = static_cast<uint32_t>(refs.size());
      pl.width = sh.width;
      pl.bits = sh.bits;
      pl.lanes = sh.lanes;
      pl.cfg = ChooseCfg(E, cfg_len[sh.cfgw == 32], sh.cfgw, true);
      pl.W = (len + sh.lanes - 1) / sh.lanes * sh.lanes;
      pl.stride = static_cast<uint32_t>(RoundUp(pl.W, static_cast<uint64_t>(pl.cfg.k)));
      pl.tiles = JobTiles(pl.stride, pl.cfg);
      if (pl.tiles > kMaxSplitTiles) {  // such long rows run their tiles in turn
        pl.cfg.split = 0;
        pl.tiles = 1;
      }
      pl.mat_bytes = JobBytes(P, V, pl.stride, sh.bits, pl.cfg);
      pl.graph_words = 0;  // the window's block is resident
      pl.seq_bytes = 0;
      pl.work = static_cast<uint64_t>(V) * len;
      if (pl.mat_bytes > E->budget) return WinFail(X, SPOA_HIP_NO_MEMORY, E->msg->no_memory);
      plans.push_back(pl);
      refs.push_back(JobRef{w, d});
    }
  }
  std::stable_sort(plans.begin(), plans.end(), [&](const Plan& a, const Plan& b) {
    if (a.bits != b.bits) return a.bits < b.bits;
    const int la = GroupLanes(P.subtype, a), lb = GroupLanes(P.subtype, b);
    if (la != lb) return la > lb;
    if (a.cfg.k != b.cfg.k) return a.cfg.k < b.cfg.k;
    if (a.cfg.waves != b.cfg.waves) return a.cfg.waves < b.cfg.waves;
    if (a.cfg.split != b.cfg.split) return a.cfg.split < b.cfg.split;
    return a.work > b.work;
  });
  const size_t nj = plans.size();
  // table: DevJob[nj] | MergeItem[live]; pair slots per job, in table order
  std::vector<uint64_t> pair_at(nj + 1, 0);
  for (size_t x = 0; x < nj; ++x) {
    const JobRef& jr = refs[plans[x].job];
    pair_at[x + 1] = pair_at[x] + X->state[jr.window].V + X->read_len[X->read_first[jr.window] + r];
  }
  if (pair_at[nj] > 0xffffffffull) return WinFail(X, SPOA_HIP_NO_MEMORY, E->msg->no_memory);
  const auto t_alloc0 = std::chrono::steady_clock::now();
  const uint64_t off_items = RoundUp(nj * sizeof(DevJob), kAlign);
  // queue items of split jobs (every tile an item), after the merge items
  uint64_t n_tiles = 0;
  bool any_split = false;
  for (size_t x = 0; x < nj; ++x) {
    n_tiles += plans[x].tiles;
    any_split |= plans[x].tiles > 1;
  }
  const uint64_t off_tiles = RoundUp(off_items + X->live.size() * sizeof(MergeItem), kAlign);
  const uint64_t table_bytes = RoundUp(off_tiles + (any_split ? n_tiles * sizeof(uint2) : 0), kAlign) + kAlign;
  if (!X->h_table.Reserve(table_bytes) || !X->d_table.Reserve(table_bytes, E->msg->no_memory) ||
      !X->d_res.Reserve(std::max<uint64_t>(nj, 1) * 16, E->msg->no_memory) ||
      !X->d_pairs.Reserve(std::max<uint64_t>(pair_at[nj], 1) * 8, E->msg->no_memory)) {
    return WinFail(X, SPOA_HIP_NO_MEMORY, g_last_error);
  }
  // passes: DP matrices within the budget; the arena is reused pass by pass
  std::vector<spoa_amd::PassItem> items(nj);
  for (size_t x = 0; x < nj; ++x) items[x] = spoa_amd::PassItem{plans[x].mat_bytes, 0, 0, pair_at[x + 1] - pair_at[x]};
  std::vector<std::pair<size_t, size_t>> passes;
  for (size_t lo = 0; lo < nj;) {
    const size_t hi = spoa_amd::PassEnd(items.data(), lo, nj, E->budget);
    passes.push_back({lo, hi});
    lo = hi;
  }
  // Every pass places its matrices from the start of one arena segment
  // large enough for the largest pass (passes run one after another on the
  // stream).  The placements of all passes are made before any of them
  // launches, so no placement may free a segment: when none is large
  // enough, the arena is rebuilt here, while the stream is idle (the
  // previous round has been waited for).
  {
    uint64_t largest_pass = 0;
    for (const auto& ps : passes) {
      uint64_t b = 0;
      for (size_t x = ps.first; x < ps.second; ++x) b += plans[x].mat_bytes;
      largest_pass = std::max(largest_pass, b);
    }
    uint64_t largest_seg = 0;
    for (const auto& sg : E->arena) largest_seg = std::max(largest_seg, sg.size);
    if (largest_pass > largest_seg) {
      if (E->pending.active || !HipOk(hipStreamSynchronize(E->stream), "hipStreamSynchronize")) {
        return WinFail(X, SPOA_HIP_DEVICE_ERROR, "[spoa::hip] the engine's stream is busy");
      }
      for (auto& sg : E->arena) sg.Release();
      E->arena.clear();
      E->arena_bytes = 0;
      DevBuf sg;
      uint64_t want = std::min<uint64_t>(E->budget, std::max<uint64_t>(2 * largest_pass, 4ull << 30));
      if (hipMalloc(&sg.p, want) != hipSuccess) {
        (void)hipGetLastError();
        want = largest_pass;
        if (hipMalloc(&sg.p, want) != hipSuccess) {
          (void)hipGetLastError();
          return WinFail(X, SPOA_HIP_NO_MEMORY, E->msg->no_memory);
        }
      }
      sg.size = want;
      E->arena.push_back(sg);
      E->arena_bytes = want;
    }
  }
  // hand-off words of split jobs: the passes run one after another on the
  // stream, so each pass's words start over (and every launch has its own
  // epoch); sized for the largest pass before anything launches
  std::vector<uint64_t> x_at(nj, 0);
  if (any_split) {
    uint64_t most = 0;
    for (const auto& ps : passes) {
      uint64_t w = spoa_hip_dev::kXferFirst;
      for (size_t x = ps.first; x < ps.second; ++x) {
        if (plans[x].tiles > 1) {
          x_at[x] = w;
          w += XferWords(X->state[refs[plans[x].job].window].V, plans[x].tiles);
        }
      }
      most = std::max(most, w);
    }
    if (most > 0xffffffffull || !ReserveXfer(E, most)) return WinFail(X, SPOA_HIP_NO_MEMORY, E->msg->no_memory);
    uint2* it = X->h_table.At<uint2>(off_tiles);
    for (const auto& ps : passes) {
      size_t g0 = ps.first;
      for (size_t x = ps.first; x < ps.second; ++x) {
        if (x > ps.first && !SameGroup(P.subtype, plans[x], plans[x - 1])) g0 = x;
        for (uint32_t t = 0; t < plans[x].tiles; ++t) *it++ = make_uint2(static_cast<uint32_t>(x - g0), t);
      }
    }
  }
  DevJob* dj = X->h_table.At<DevJob>(0);
  std::vector<int32_t> job_of(refs.size(), -1);  // ref -> table index
  std::vector<std::vector<uint64_t>> pass_addr(passes.size());
  for (size_t pi = 0; pi < passes.size(); ++pi) {
    const size_t lo = passes[pi].first, hi = passes[pi].second;
    std::vector<uint64_t> need(hi - lo);
    for (size_t x = lo; x < hi; ++x) need[x - lo] = plans[x].mat_bytes;
    const int rc = PlaceMatrices(E, need, &pass_addr[pi], kBatchSegment);
    if (rc != SPOA_HIP_OK) return WinFail(X, rc, g_last_error);
    for (size_t x = lo; x < hi; ++x) {
      const Plan& pl = plans[x];
      const JobRef& jr = refs[pl.job];
      const uint32_t w = jr.window;
      const uint32_t k = X->read_first[w] + r;
      DevJob& d = dj[x];
      d = DevJob{};
      d.mat_addr = pass_addr[pi][x - lo];
      d.graph_off = X->info[w].block_off;
      d.seq_off = jr.strand ? X->seq_r[k] : X->seq_f[k];
      d.V = X->state[w].V;
      d.n = X->read_len[k];
      d.W = pl.W;
      d.stride = pl.stride;
      d.pair_off = static_cast<uint32_t>(pair_at[x]);
      d.cap = d.V + d.n;
      d.nslots = X->state[w].nslots;
      d.slot_chars = X->state[w].slot_chars;
      d.split = pl.tiles > 1 ? pl.tiles : 0u;
      d.xfer_off = static_cast<uint32_t>(x_at[x]);
      job_of[pl.job] = static_cast<int32_t>(x);
      const uint64_t jb = static_cast<uint64_t>(pl.bits / 8) * d.n *
                          (static_cast<uint64_t>(Matrices(P.subtype)) * d.V +
                           static_cast<uint64_t>(ReadMatrices(P.subtype)) * X->state[w].R);
      X->stats.dp_bytes += jb;
      (pl.width == 16 ? X->stats.dp_bytes16 : X->stats.dp_bytes32) += jb;
      X->stats.cells += static_cast<uint64_t>(d.V) * d.n;
      X->stats.jobs += 1;
    }
  }
  X->stats.alloc_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_alloc0).count();
  // merge items (job indices by window)
  {
    std::vector<int32_t> jf(X->n, -1), jr(X->n, -1);
    for (size_t q = 0; q < refs.size(); ++q) (refs[q].strand ? jr : jf)[refs[q].window] = job_of[q];
    MergeItem* mi = X->h_table.At<MergeItem>(off_items);
    for (size_t i = 0; i < X->live.size(); ++i) {
      const uint32_t w = X->live[i];
      const uint32_t k = X->read_first[w] + r;
      mi[i] = MergeItem{w, jf[w], jr[w], r, X->read_len[k], X->seq_f[k], X->seq_r[k], X->qual[k]};
    }
  }
  X->stats.host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  hipStream_t st = E->stream;
  uint8_t* dtab = static_cast<uint8_t*>(X->d_table.p);
  if (!HipOk(hipMemcpyAsync(dtab, X->h_table.p, table_bytes, hipMemcpyHostToDevice, st), "copy in")) {
    return SPOA_HIP_DEVICE_ERROR;
  }
  // the DP, pass by pass
  X->launches = 0;
  X->launch_width.clear();
  uint64_t pass_tiles = 0;  // items of the passes before this one
  for (size_t pi = 0; pi < passes.size(); ++pi) {
    const size_t lo = passes[pi].first, hi = passes[pi].second;
    std::vector<LaunchGroup> groups;
    const int rc = BuildGroups(E, plans, lo, hi, &groups);
    if (rc != SPOA_HIP_OK) return WinFail(X, rc, g_last_error);
    if (groups.size() > 256) return WinFail(X, SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] too many launch groups");
    uint32_t* counters = static_cast<uint32_t*>(X->d_counter.p);
    if (!LaunchResetCounters(counters, static_cast<uint32_t>(groups.size()), st)) return SPOA_HIP_DEVICE_ERROR;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      const LaunchGroup& gp = groups[gi];
      while (X->ev_s.size() <= X->launches) {
        hipEvent_t a = nullptr, b = nullptr;
        if (!HipOk(hipEventCreate(&a), "hipEventCreate") || !HipOk(hipEventCreate(&b), "hipEventCreate")) {
          return SPOA_HIP_DEVICE_ERROR;
        }
        X->ev_s.push_back(a);
        X->ev_e.push_back(b);
      }
      KArgs args{};
      args.prm = gp.prm;
      args.jobs = reinterpret_cast<const DevJob*>(dtab) + lo + gp.a;
      args.n_jobs = static_cast<uint32_t>(gp.b - gp.a);
      args.queue = counters + gi;
      args.graphs = static_cast<const uint32_t*>(X->d_blocks.p);
      args.seqs = static_cast<const uint8_t*>(X->d_seqs.p);
      args.pairs = static_cast<int32_t*>(X->d_pairs.p);
      args.results = static_cast<int32_t*>(X->d_res.p) + 4 * (lo + gp.a);
      args.lds_bytes = gp.lds;
      args.prio_jobs = gp.prio;
      args.phases = E->phases;
      if (any_split) {
        args.items = reinterpret_cast<const uint2*>(dtab + off_tiles) + pass_tiles + gp.item_at;
        args.n_items = gp.items;
      }
      args.epoch = NextEpoch(E);
      args.wait_ticks = E->wait_ticks;
      args.xfer = static_cast<unsigned long long*>(E->d_xfer.p);
      if (!HipOk(hipEventRecord(X->ev_s[X->launches], st), "event")) return SPOA_HIP_DEVICE_ERROR;
      gp.fn<<<gp.grid, gp.threads, gp.lds, st>>>(args);
      if (!HipOk(hipGetLastError(), "kernel launch") || !HipOk(hipEventRecord(X->ev_e[X->launches], st), "event")) {
        return SPOA_HIP_DEVICE_ERROR;
      }
      X->launch_width.push_back(plans[lo + gp.a].width);
      ++X->launches;
    }
    for (size_t x = lo; x < hi; ++x) pass_tiles += plans[x].tiles;
  }
  // the merge
  if (!X->live.empty()) {
    spoa_hip_dev::MergeArgs ma{};
    ma.items = reinterpret_cast<const MergeItem*>(dtab + off_items);
    ma.info = static_cast<const WinInfo*>(X->d_info.p);
    ma.windows = static_cast<uint8_t*>(X->d_windows.p);
    ma.blocks = static_cast<uint32_t*>(X->d_blocks.p);
    ma.jobs = reinterpret_cast<const DevJob*>(dtab);
    ma.results = static_cast<const int32_t*>(X->d_res.p);
    ma.pairs = static_cast<const int32_t*>(X->d_pairs.p);
    ma.seqs = static_cast<const uint8_t*>(X->d_seqs.p);
    ma.quals = X->quals ? static_cast<const uint8_t*>(X->d_quals.p) : nullptr;
    ma.summary = static_cast<WinSummary*>(X->d_summary.p);
    ma.type = P.type;
    ma.subtype = P.subtype;
    ma.g = P.g;
    ma.e = P.e;
    ma.q = P.q;
    ma.c = P.c;
    ma.lds_bytes = MergeLds(X, r, P.type == SPOA_HIP_NW);
    if (!HipOk(hipEventRecord(X->ev_m0, st), "event")) return SPOA_HIP_DEVICE_ERROR;
    spoa_hip_dev::MergeKernel<<<static_cast<uint32_t>(X->live.size()), 64, ma.lds_bytes, st>>>(ma);
    if (!HipOk(hipGetLastError(), "kernel launch") || !HipOk(hipEventRecord(X->ev_m1, st), "event")) {
      return SPOA_HIP_DEVICE_ERROR;
    }
  }
  if (!HipOk(hipMemcpyAsync(X->h_summary.p, X->d_summary.p, X->n * sizeof(WinSummary), hipMemcpyDeviceToHost, st),
             "copy out") ||
      !HipOk(hipEventRecord(X->ev_done, st), "event")) {
    return SPOA_HIP_DEVICE_ERROR;
  }
  X->stats.launches += static_cast<uint32_t>(X->launches);
  X->pending = true;
  return SPOA_HIP_OK;
}

int spoa_hip_windows_ready(spoa_hip_windows* X) {
  if (!X) return -Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null argument");
  if (!X->pending) return 1;
  const hipError_t q = hipEventQuery(X->ev_done);
  if (q == hipErrorNotReady) return 0;
  (void)hipGetLastError();
  return 1;
}

int spoa_hip_windows_wait(spoa_hip_windows* X) {
  if (!X) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null argument");
  if (!X->pending) return SPOA_HIP_OK;
  X->pending = false;
  const auto t0 = std::chrono::steady_clock::now();
  if (!HipOk(hipSetDevice(X->E->device), "hipSetDevice") || !HipOk(hipEventSynchronize(X->ev_done), "round")) {
    return SPOA_HIP_DEVICE_ERROR;
  }
  X->stats.wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (size_t i = 0; i < X->launches; ++i) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, X->ev_s[i], X->ev_e[i]);
    X->stats.kernel_ms += ms;
    if (X->launch_width[i] == 16) {
      X->stats.kernel_ms16 += ms;
      X->stats.launches16 += 1;
    } else {
      X->stats.kernel_ms32 += ms;
      X->stats.launches32 += 1;
    }
  }
  if (!X->live.empty()) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, X->ev_m0, X->ev_m1);
    X->stats.merge_ms += ms;
    X->stats.merges += static_cast<uint32_t>(X->live.size());
  }
  const WinSummary* s = X->h_summary.At<WinSummary>(0);
  int rc = SPOA_HIP_OK;
  for (uint32_t w : X->live) {
    X->state[w] = s[w];
    if (s[w].status != wg::kWinOk && rc == SPOA_HIP_OK) {
      rc = s[w].status == wg::kWinCapacity ? SPOA_HIP_NO_MEMORY : SPOA_HIP_INCONSISTENT;
      WinFail(X, rc, std::string(WinStatusText(s[w].status)) + " (window " + std::to_string(w) + ")");
    }
  }
  return rc;
}

int spoa_hip_windows_stats(spoa_hip_windows* X, spoa_hip_stats* out) {
  if (!X || !out) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null argument");
  *out = X->stats;
  return SPOA_HIP_OK;
}

int spoa_hip_windows_finish(spoa_hip_windows* X, int32_t msa) {
  if (!X) return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] null argument");
  if (X->pending) {
    const int rc = spoa_hip_windows_wait(X);
    if (rc != SPOA_HIP_OK) return rc;
  }
  spoa_hip_engine* E = X->E;
  if (!HipOk(hipSetDevice(E->device), "hipSetDevice")) return SPOA_HIP_DEVICE_ERROR;
  hipStream_t st = E->stream;
  const uint32_t n = X->n;
  if (n == 0) {
    X->finished = true;
    return SPOA_HIP_OK;
  }
  auto* info = static_cast<const WinInfo*>(X->d_info.p);
  auto* win = static_cast<uint8_t*>(X->d_windows.p);
  auto* blk = static_cast<uint32_t*>(X->d_blocks.p);
  auto* sum = static_cast<WinSummary*>(X->d_summary.p);
  spoa_hip_dev::ConsensusKernel<<<n, 64, 0, st>>>(info, win, blk, sum, n);
  if (!HipOk(hipGetLastError(), "kernel launch") ||
      !HipOk(hipMemcpyAsync(X->h_summary.p, X->d_summary.p, n * sizeof(WinSummary), hipMemcpyDeviceToHost, st),
             "copy out") ||
      !HipOk(hipStreamSynchronize(st), "consensus")) {
    return SPOA_HIP_DEVICE_ERROR;
  }
  const WinSummary* s = X->h_summary.At<WinSummary>(0);
  X->cons_off.assign(n + 1, 0);
  X->msa_off.assign(n + 1, 0);
  X->width.assign(n, 0);
  X->nseq.assign(n, 0);
  uint64_t total = 0;
  for (uint32_t w = 0; w < n; ++w) {
    if (s[w].status != wg::kWinOk) {
      return WinFail(X, SPOA_HIP_INCONSISTENT, std::string(WinStatusText(s[w].status)) + " (window " +
                                                    std::to_string(w) + ")");
    }
    X->cons_off[w] = total;
    total += s[w].aux;
    X->cons_off[w + 1] = total;
  }
  const uint64_t cons_total = total;
  for (uint32_t w = 0; w < n; ++w) {
    X->width[w] = s[w].A;
    X->nseq[w] = s[w].E;
    X->msa_off[w] = total;
    if (msa) total += static_cast<uint64_t>(s[w].E) * s[w].A;
    X->msa_off[w + 1] = total;
  }
  (void)cons_total;
  const uint64_t offs_bytes = 2 * (n + 1) * sizeof(uint64_t);
  if (!X->h_offs.Reserve(offs_bytes) || !X->d_offs.Reserve(offs_bytes, E->msg->no_memory) ||
      !X->d_out.Reserve(std::max<uint64_t>(total, 256), E->msg->no_memory) ||
      !X->h_out.Reserve(std::max<uint64_t>(total, 256))) {
    return WinFail(X, SPOA_HIP_NO_MEMORY, g_last_error);
  }
  std::memcpy(X->h_offs.At<uint64_t>(0), X->cons_off.data(), (n + 1) * sizeof(uint64_t));
  std::memcpy(X->h_offs.At<uint64_t>((n + 1) * sizeof(uint64_t)), X->msa_off.data(), (n + 1) * sizeof(uint64_t));
  const uint64_t* d_co = static_cast<const uint64_t*>(X->d_offs.p);
  if (!HipOk(hipMemcpyAsync(X->d_offs.p, X->h_offs.p, offs_bytes, hipMemcpyHostToDevice, st), "copy in")) {
    return SPOA_HIP_DEVICE_ERROR;
  }
  spoa_hip_dev::OutputKernel<<<n, 64, 0, st>>>(info, win, blk, d_co, d_co + (n + 1), msa,
                                               static_cast<uint8_t*>(X->d_out.p), n);
  if (!HipOk(hipGetLastError(), "kernel launch") ||
      !HipOk(hipMemcpyAsync(X->h_out.p, X->d_out.p, total, hipMemcpyDeviceToHost, st), "copy out") ||
      !HipOk(hipStreamSynchronize(st), "output")) {
    return SPOA_HIP_DEVICE_ERROR;
  }
  X->finished = true;
  X->msa_done = msa != 0;
  return SPOA_HIP_OK;
}

int spoa_hip_windows_consensus(spoa_hip_windows* X, uint32_t w, const char** out, uint32_t* len) {
  if (!X || !out || !len || !X->finished || w >= X->n) {
    return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] no such finished window");
  }
  *out = X->h_out.At<char>(X->cons_off[w]);
  *len = static_cast<uint32_t>(X->cons_off[w + 1] - X->cons_off[w]);
  return SPOA_HIP_OK;
}

int spoa_hip_windows_msa(spoa_hip_windows* X, uint32_t w, const char** rows, uint32_t* n_rows, uint32_t* width) {
  if (!X || !rows || !n_rows || !width || !X->finished || w >= X->n) {
    return Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] no such finished window");
  }
  *rows = X->h_out.At<char>(X->msa_off[w]);
  *n_rows = X->msa_done ? X->nseq[w] : 0;
  *width = X->msa_done ? X->width[w] : 0;
  return SPOA_HIP_OK;
}

int spoa_hip_windows_reversed(spoa_hip_windows* X, uint32_t w, uint8_t* out, uint32_t cap) {
  if (!X || w >= X->n) return -Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] no such window");
  const uint32_t n = X->nreads[w];
  if (n > cap || (!out && n)) return -Fail(SPOA_HIP_CAPACITY, "[spoa::hip] buffer too small");
  if (!n) return 0;
  if (!HipOk(hipSetDevice(X->E->device), "hipSetDevice")) return -SPOA_HIP_DEVICE_ERROR;
  const wg::WinLayout l = wg::MakeLayout(X->info[w].caps);
  if (!HipOk(hipMemcpy(out, static_cast<uint8_t*>(X->d_windows.p) + X->info[w].region + l.reversed, n,
                       hipMemcpyDeviceToHost),
             "copy out")) {
    return -SPOA_HIP_DEVICE_ERROR;
  }
  return static_cast<int>(n);
}

int spoa_hip_windows_ranks(spoa_hip_windows* X, uint32_t w, uint32_t* out, uint32_t cap) {
  if (!X || w >= X->n) return -Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] no such window");
  if (X->pending) return -Fail(SPOA_HIP_INVALID_ARGUMENT, "[spoa::hip] a round is in flight");
  const uint32_t V = X->state[w].V;
  if (V > cap || (!out && V)) return -Fail(SPOA_HIP_CAPACITY, "[spoa::hip] buffer too small");
  if (!V) return 0;
  if (!HipOk(hipSetDevice(X->E->device), "hipSetDevice")) return -SPOA_HIP_DEVICE_ERROR;
  const wg::WinLayout l = wg::MakeLayout(X->info[w].caps);
  if (!HipOk(hipMemcpy(out, static_cast<uint8_t*>(X->d_windows.p) + X->info[w].region + l.rank_to_node,
                       4ull * V, hipMemcpyDeviceToHost),
             "copy out")) {
    return -SPOA_HIP_DEVICE_ERROR;
  }
  return static_cast<int>(V);
}

const char* spoa_hip_windows_last_error(spoa_hip_windows* X) { return X ? X->last_error.c_str() : ""; }

}  // extern "C"
