// Device side of the graph-to-sequence alignment: the DP over a
// rank-ordered POA graph and the on-device traceback, for CDNA4 (gfx950).
//
// Semantics follow the reference SIMD engine
// (src/simd_alignment_engine_implementation.hpp of spoa v4.0.8, AVX2 lane
// counts): rows of W = ceil(n / L) * L columns with L = 16 (its int16
// engine) or L = 8 (its int32 engine; the kernels take L as a template
// parameter, so int16 rows can reproduce the int32 engine), padding
// columns scored with
// -max(|m|, |n|, |g|, |q|) (:526-542), linear / affine / convex
// recurrences (:824-901, :1179-1272, :1638-1774), the end-cell choice
// (:903-953) and the traceback with its tie order (:955-1124, :1324-1574,
// :1826-2121).  In convex mode a gap in E/Q opens from the pre-gap score
// H1 inside an L-column segment and from the final score at a segment
// start (:1738-1768); the kernel keeps that segment rule exactly.
//
// Work decomposition.  One workgroup aligns one (graph, read) job at a
// time and takes jobs from a queue (a persistent grid).  Lanes own K
// consecutive columns (K a multiple of L in convex mode, so every lane
// starts a segment); a row is one tile of nwaves * 64 * K columns or a
// few of them.  Each row runs
//   1. the vertical pass: every predecessor row (H, F, O) is read back
//      with 16-byte loads; the diagonal H_p[c0 - 1] comes from the
//      neighbouring lane by a cross-lane shuffle;
//   2. the horizontal pass as a max-plus scan: every lane summarises its K
//      columns as a transfer Y_out = M (x) Y_in (+) b, where Y is the
//      candidate value entering the lane's first column (H for linear, E
//      for affine, the pair (E, Q) for convex).  M is the same for every
//      lane (a scalar decay K*g or K*e, or a 2x2 max-plus matrix), so the
//      wavefront scan only needs its powers M^(2^t); waves are chained
//      through LDS;
//   3. a second pass over the K columns with the exact carry, the stores
//      of the H/F/E/O/Q rows and the end-cell bookkeeping.
// No sentinel ever takes part in the arithmetic: aggregates are formed
// with the carry term absent, so int32 scores may use their full range.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "poa_types.hpp"

namespace spoa_hip_dev {

// ------------------------------------------------------------- helpers --

// y <- M (x) y for a 2x2 max-plus matrix (row-major).
__device__ __forceinline__ void MpApply(const int32_t* M, int32_t& y0, int32_t& y1) {
  const int32_t a = max(M[0] + y0, M[1] + y1);
  const int32_t b = max(M[2] + y0, M[3] + y1);
  y0 = a;
  y1 = b;
}

// out <- M^k (k >= 1) from the powers M^(2^t) (max-plus; powers of one
// matrix commute).
__device__ __forceinline__ void LanePow(const KParams& prm, int k, int32_t (&out)[4]) {
  bool have = false;
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    if ((k >> t) & 1) {
      const int32_t* b = prm.mpow[t];
      if (!have) {
        out[0] = b[0];
        out[1] = b[1];
        out[2] = b[2];
        out[3] = b[3];
        have = true;
      } else {
        const int32_t r0 = max(out[0] + b[0], out[1] + b[2]);
        const int32_t r1 = max(out[0] + b[1], out[1] + b[3]);
        const int32_t r2 = max(out[2] + b[0], out[3] + b[2]);
        const int32_t r3 = max(out[2] + b[1], out[3] + b[3]);
        out[0] = r0;
        out[1] = r1;
        out[2] = r2;
        out[3] = r3;
      }
    }
  }
}

// DPP lane moves (gfx9 encodings): row_shr:n within rows of 16 lanes,
// wave_shr:1 across the wave.  Lanes whose source is out of range keep old.
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kWaveShr1 = 0x138;

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int32_t Dpp(int32_t old, int32_t v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, 0xf, false);
}

// Waits for every vector memory operation of the wave (vmcnt 0), through
// the builtin so that the compiler knows the loads before it have landed.
// Vector loads and stores share one in-order counter: a load issued before
// a row's stores and waited for after them would wait for the stores too,
// so every load is waited for before the stores that follow it.
__device__ __forceinline__ void WaitVmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Barrier for LDS traffic only: waits for this wave's LDS operations, not
// for its global loads and stores (which __syncthreads would also drain).
__device__ __forceinline__ void LdsBarrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Row-phase timing (a build with SPOA_ROW_PHASES only): wave 0 of every
// workgroup reads the shader clock (s_memtime) at the phase boundaries of
// each row tile and adds the cycles of each phase into KArgs::phases at
// the end of each job.  Phases (the RowPhase order): query profile,
// vertical pass (of which: waiting for predecessor loads), local pass,
// scan (of which: the wave-total barrier), fix-up and stores (of which:
// the wait before the first store), end cell and boundary hand-off (of
// which: its barrier); then per job: end-cell reduction and traceback;
// and counts of row tiles and jobs.
enum RowPhase {
  kPhProfile = 0, kPhVertical, kPhPredWait, kPhLocal, kPhScan, kPhScanBarrier, kPhFixup, kPhStoreWait,
  kPhEndBoundary, kPhBoundaryBarrier, kPhReduce, kPhTraceback, kPhTiles, kPhJobs, kPhCount
};
// Traceback parts, after the row phases: tile staging (cycles, count),
// diagonal runs (cycles, steps), general steps (cycles, count), left and up
// gap walks (cycles).
enum TbPhase { kTbStage = 0, kTbStages, kTbDiag, kTbDiagSteps, kTbGeneral, kTbGenerals, kTbLeft, kTbUp, kTbCount };
constexpr int kPhTotal = kPhCount + kTbCount;
#ifdef SPOA_ROW_PHASES
// the scheduler may not move work across a timestamp
#define SPOA_PH_NOW()                                    \
  ({                                                     \
    __builtin_amdgcn_sched_barrier(0);                   \
    unsigned long long t_ = __builtin_amdgcn_s_memtime(); \
    __builtin_amdgcn_sched_barrier(0);                   \
    t_;                                                  \
  })
#define SPOA_PH_ADD(k, v) (ph[k] += (v))
#define SPOA_TB_ADD(k, v) (tph[k] += (v))
#else
#define SPOA_PH_NOW() 0ull
#define SPOA_PH_ADD(k, v) ((void)0)
#define SPOA_TB_ADD(k, v) ((void)0)
#endif

__device__ __forceinline__ uint32_t Uniform(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

// ------------------------------------------------- split-job hand-off --
// The row tiles of a split job run on different workgroups (possibly on
// different XCDs, whose L2 caches are not coherent with each other).  They
// exchange 64-bit words: the launch's tag in the high half, a value in the
// low half, stored and loaded with device-scope relaxed atomics (vector
// accesses that go to the device's coherence point).  A word is valid once
// it carries the launch's tag, so no fence orders the per-row hand-off;
// only the traceback, which reads every tile's rows, pairs a release fence
// of each tile with an acquire fence of the tile that walks it.
__device__ __forceinline__ void XferPut(unsigned long long* p, uint32_t tag, int32_t v) {
  __hip_atomic_store(p, (static_cast<unsigned long long>(tag) << 32) | static_cast<uint32_t>(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long XferGet(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool XferTagged(unsigned long long w, uint32_t tag) {
  return static_cast<uint32_t>(w >> 32) == tag;
}

// Waits (the whole wave) until the word that each lane l < nw holds in w --
// loaded from base[l] -- carries the launch's tag, reloading it after a
// short sleep.  Gives up after wait_ticks of the 100 MHz clock, raising the
// launch's abort word xfer[0], or at once when that word is already raised.
// True when every word is valid.
__device__ __forceinline__ bool XferWait(unsigned long long* xfer, uint32_t tag, uint32_t wait_ticks,
                                         unsigned long long* base, int nw, int lane, unsigned long long& w) {
  auto pending = [&]() { return __builtin_amdgcn_ballot_w64(lane < nw && !XferTagged(w, tag)) != 0; };
  if (!pending()) return true;
  const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
  for (;;) {
    const unsigned long long ab = XferGet(xfer);
    if (__builtin_amdgcn_ballot_w64(XferTagged(ab, tag)) != 0) return false;
    __builtin_amdgcn_s_sleep(1);
    if (lane < nw) w = XferGet(base + lane);
    if (!pending()) return true;
    if (__builtin_amdgcn_s_memrealtime() - t0 > wait_ticks) {
      if (lane == 0) XferPut(xfer, tag, 1);
      return false;
    }
  }
}

// The K read bases at columns c0.. (4 per word) and how many of them lie
// inside the read; the sequence block is padded to the row stride.
template <int K>
__device__ __forceinline__ void LoadChars(const uint8_t* __restrict__ seq, uint32_t c0, uint32_t n, bool active,
                                          uint32_t (&scp)[K / 4], int32_t& nvalid) {
  if (active) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(seq + c0);
#pragma unroll
    for (int i = 0; i < K / 4; ++i) scp[i] = w[i];
  } else {
#pragma unroll
    for (int i = 0; i < K / 4; ++i) scp[i] = 0u;
  }
  const int32_t left = static_cast<int32_t>(n) - static_cast<int32_t>(c0);
  nvalid = left < 0 ? 0 : (left > K ? K : left);
}

// ------------------------------------------------------------ traceback --
// The reference traceback, walked by one wave.  Every lane runs the same
// serial walk (the same values, so control flow stays uniform); the lanes
// together stage an LDS tile of what the walk reads next -- rows [r0, r1]
// and columns [c0, c0 + C) of every stored matrix, the read's bases at those
// columns, the predecessor offsets, node ids and bases of those rows and as
// many of their predecessor lists as fit -- so that a step reads LDS instead of a chain of dependent device
// loads.  The walk only moves to lower rows and columns; when it nears the
// tile's lower or left edge it restages the tile below and left of its
// position.  Reads outside the tile (long edges) go to device memory.
// Pairs are emitted in traceback order (the host reverses them); lane 0
// stores them.
template <typename T>
using Lds = __attribute__((address_space(3))) T;
// Device-memory reads of the walk outside its tile (long edges), kept out
// of line so that the tile reads stay LDS reads.
template <typename T>
__device__ __attribute__((noinline)) T TbLoad(const T* p) {
  return *p;
}

template <typename ST>
struct TbTile {
  Lds<ST>* m;         // [R][NM][C]
  Lds<uint32_t>* po;  // [R + 1] predecessor offsets of rows rb..r1 (and the end of r1)
  Lds<U32x4>* rec;    // [R] per row: first predecessor (0 for a source), base, node id
  Lds<uint32_t>* pr;  // [P] predecessors pb0 .. pb0 + pn
  Lds<uint8_t>* sq;   // [C] read bases of the tile columns
  int32_t R = 0, C = 0, cshift = 0, P = 0, margin = 0;
  bool on = false;
  int32_t r0 = 1, r1 = 0, rb = 1, c0 = 0;
  uint32_t pb0 = 0, pn = 0;
};

template <int SUB, typename ST, bool SISD>
__device__ void Traceback(const KParams& prm, const GraphView& gv, const uint8_t* __restrict__ seq,
                          uint32_t n, const ST* __restrict__ mat, uint64_t rstride, uint32_t stride,
                          int32_t max_i, int32_t max_j, int32_t* __restrict__ out, uint32_t cap,
                          uint32_t* n_out, int32_t* status_out, Lds<uint32_t>* lds, uint32_t lds_bytes, int lane,
                          unsigned long long (&tph)[kTbCount]) {
  (void)tph;
  constexpr int NM = SUB == 0 ? 1 : (SUB == 1 ? 3 : 5);
  constexpr int S = sizeof(ST);
  constexpr int VEC = 16 / S;  // columns a 16-byte load
  const int32_t g = prm.g, e = prm.e, q = prm.q, c = prm.c;
  const int type = prm.type;
  // The scalar engine's walk (sisd...:361-437, :547-676, :777-923): the
  // same tests in the same order, but its global walk runs on through row 0
  // and the first column (column -1 here) in the main loop, its convex left
  // walk goes on while either gap chain holds, and E / Q of the first
  // column are -inf (0 in row 0).
  constexpr bool sisd = SISD;
  constexpr int32_t kNegInf = -2147483647 - 1 + 1024;
  auto U = [](int32_t v) -> int32_t { return static_cast<int32_t>(Uniform(static_cast<uint32_t>(v))); };

  // Tile geometry from the LDS the kernel lends (its query profile, dead
  // now): a power-of-two column count C and as many rows as fit, chosen so
  // that the walk -- about max_i / max_j rows a column -- runs longest
  // before it leaves the tile.
  TbTile<ST> T;
  {
    const uint32_t avail = lds_bytes > 128 ? lds_bytes - 128 : 0;  // spare: end offset, bases
    const int32_t ratio16 = min(64, max(16, static_cast<int32_t>((16ll * max(max_i, 1)) / (max_j + 1))));
    int32_t best = -1;
    for (int32_t cs = 4; cs <= 6; ++cs) {  // C = 16, 32, 64
      const int32_t C = 1 << cs;
      if (C < 2 * VEC || C > static_cast<int32_t>(stride)) continue;
      // matrices, offset, record, ~3 predecessors a row
      const int32_t R = static_cast<int32_t>(avail / static_cast<uint32_t>(NM * C * S + 4 + 16 + 3 * 4));
      if (R < 8) continue;
      const int32_t steps = min(C - VEC - 2, ((R - 4) * 16) / ratio16);
      if (steps > best) {
        best = steps;
        T.R = R;
        T.C = C;
        T.cshift = cs;
      }
    }
    if (best > 0) {
      const int32_t R = T.R;
      T.P = 3 * R;
      T.margin = 4;
      T.m = reinterpret_cast<Lds<ST>*>(lds);
      T.rec = reinterpret_cast<Lds<U32x4>*>(lds + (NM * R * T.C * S) / 4);
      T.po = reinterpret_cast<Lds<uint32_t>*>(T.rec + R);
      T.pr = T.po + R + 1;
      T.sq = reinterpret_cast<Lds<uint8_t>*>(T.pr + T.P);
      T.on = true;
    }
  }
  // stage the tile whose top-right cell is (i, j + 1)
  auto stage = [&](int32_t i, int32_t j) {
    const unsigned long long t0 = SPOA_PH_NOW();
    (void)t0;
    T.r1 = i;
    T.r0 = max(0, i - T.R + 1);
    int32_t c0 = j + 2 - T.C;
    c0 = c0 <= 0 ? 0 : ((c0 + VEC - 1) / VEC) * VEC;
    if (c0 + T.C > static_cast<int32_t>(stride)) c0 = static_cast<int32_t>(stride) - T.C;
    T.c0 = c0;
    const int32_t nrow = T.r1 - T.r0 + 1;
    // 16-byte pieces of one row of one matrix: C / VEC = 1 << kps
    const int32_t kps = T.cshift - (S == 2 ? 3 : 2);
    const int32_t pieces = (nrow * NM) << kps;
    // global -> LDS loads: lane l of a round writes 16 bytes at the round's
    // LDS base + 16 l, so the tile is laid out in piece order: [row][matrix][column].
    // Nothing is pending when the rounds start (a wait the compiler sees),
    // so they issue back to back.
    __builtin_amdgcn_s_waitcnt(0);
    for (int32_t x0 = 0; x0 < pieces; x0 += 64) {
      const int32_t x = x0 + lane;
      if (x < pieces) {
        const int32_t k = x & ((1 << kps) - 1), rt = x >> kps, t = rt % NM, r = rt / NM;
        const ST* src = mat + static_cast<uint64_t>(T.r0 + r) * rstride + static_cast<uint64_t>(t) * stride +
                        static_cast<uint32_t>(c0 + k * VEC);
        __builtin_amdgcn_global_load_lds(static_cast<const void*>(src),
                                         reinterpret_cast<Lds<void>*>(T.m + x0 * VEC), 16, 0, 0);
      }
    }
    // rows rb..r1 carry graph data (row 0 has none)
    T.rb = max(T.r0, 1);
    const int32_t nm = T.r1 - T.rb + 1;
    for (int32_t x = lane; x <= nm; x += 64) {
      T.po[x] = gv.pred_off[T.rb - 1 + x];
      if (x < nm) {
        // the row record of the view holds the base and the first predecessor
        const uint4 rw = LoadRow(gv, T.rb - 1 + x);
        T.rec[x] = U32x4{rw.y, rw.x & 0xffu, static_cast<uint32_t>(gv.node_id[T.rb - 1 + x]), 0u};
      }
    }
    T.pb0 = nm > 0 ? Uniform(gv.pred_off[T.rb - 1]) : 0u;
    const uint32_t tot = nm > 0 ? Uniform(gv.pred_off[T.r1]) - T.pb0 : 0u;
    T.pn = min(tot, static_cast<uint32_t>(T.P));
    for (uint32_t x = static_cast<uint32_t>(lane); x < T.pn; x += 64) T.pr[x] = gv.preds[T.pb0 + x];
    if (lane < T.C) T.sq[lane] = c0 + lane < static_cast<int32_t>(n) ? seq[c0 + lane] : 0;
    // the wave's LDS writes land before any lane reads them
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    SPOA_TB_ADD(kTbStage, SPOA_PH_NOW() - t0);
    SPOA_TB_ADD(kTbStages, 1);
  };
  auto ensure = [&](int32_t i, int32_t j) {
    if (!T.on) return;
    const bool inside = i <= T.r1 && (i - T.margin >= T.r0 || T.r0 == 0) && j + 1 < T.c0 + T.C &&
                        (j - 1 >= T.c0 || T.c0 == 0);
    if (!inside) stage(i, j);
  };

  // Every value the walk reads is the same in all lanes: it is moved to a
  // scalar register, so the walk's arithmetic and branches are scalar.
  // matrix t of row r at column col (t: 0 H, 1 F, 2 E, 3 O, 4 Q)
  auto at = [&](int t, int32_t r, int32_t col) -> int32_t {
    if (r >= T.r0 && r <= T.r1 && col >= T.c0 && col < T.c0 + T.C) {
      return U(static_cast<int32_t>(T.m[(((r - T.r0) * NM + t) << T.cshift) + (col - T.c0)]));
    }
    return U(static_cast<int32_t>(TbLoad(mat + static_cast<uint64_t>(r) * rstride + static_cast<uint64_t>(t) * stride +
                                         static_cast<uint32_t>(col))));
  };
  // graph data of DP row r >= 1 (view index r - 1)
  auto pred_begin = [&](int32_t r) -> uint32_t {
    return Uniform(r >= T.rb && r <= T.r1 ? T.po[r - T.rb] : TbLoad(gv.pred_off + (r - 1)));
  };
  auto pred_end = [&](int32_t r) -> uint32_t {
    return Uniform(r >= T.rb && r <= T.r1 ? T.po[r - T.rb + 1] : TbLoad(gv.pred_off + r));
  };

  // the same reads for a value that differs by lane
  auto at_v = [&](int t, int32_t r, int32_t col) -> int32_t {
    if (r >= T.r0 && r <= T.r1 && col >= T.c0 && col < T.c0 + T.C) {
      return static_cast<int32_t>(T.m[(((r - T.r0) * NM + t) << T.cshift) + (col - T.c0)]);
    }
    return static_cast<int32_t>(
        mat[static_cast<uint64_t>(r) * rstride + static_cast<uint64_t>(t) * stride + static_cast<uint32_t>(col)]);
  };
  // matrix t of row r at column col >= -1: column -1 is the first column
  // (H, F, O from the graph block; E, Q -inf there, 0 in row 0).  Only the
  // scalar walk reaches column -1 and row 0 inside its main loop (the SIMD
  // walk stops there), so the SIMD walk reads the matrices directly.
  auto cell = [&](int t, int32_t r, int32_t col) -> int32_t {
    if (!SISD || col >= 0) return at(t, r, col);
    if (t == 0) return U(gv.fcH[r]);
    if (t == 1) return U(gv.fcF[r]);
    if (t == 3) return U(gv.fcO[r]);
    return r == 0 ? 0 : kNegInf;
  };
  auto cell_v = [&](int t, int32_t r, int32_t col) -> int32_t {
    if (!SISD || col >= 0) return at_v(t, r, col);
    if (t == 0) return gv.fcH[r];
    if (t == 1) return gv.fcF[r];
    if (t == 3) return gv.fcO[r];
    return r == 0 ? 0 : kNegInf;
  };
  auto pred_v = [&](uint32_t k) -> int32_t {
    return static_cast<int32_t>(k - T.pb0 < T.pn ? T.pr[k - T.pb0] : gv.preds[k]);
  };
  auto node_of = [&](int32_t r) -> int32_t {
    return U(r >= T.rb && r <= T.r1 ? static_cast<int32_t>(T.rec[r - T.rb].z) : TbLoad(gv.node_id + (r - 1)));
  };
  auto base_of = [&](int32_t r) -> uint32_t {
    return Uniform(r >= T.rb && r <= T.r1 ? T.rec[r - T.rb].y : (LoadRow(gv, r - 1).x & 0xffu));
  };
  auto prof = [&](int32_t r, int32_t col) -> int32_t {
    if (col >= static_cast<int32_t>(n)) return prm.pad;
    const uint32_t b = Uniform(col >= T.c0 && col < T.c0 + T.C ? T.sq[col - T.c0] : TbLoad(seq + col));
    return b == base_of(r) ? prm.m : prm.mis;
  };

  uint32_t np = 0;
  int32_t status = kDevOk;
  // pair k waits in lane k % 64 until 64 pairs are gathered, then the wave
  // stores them in one go
  int32_t buf_a = 0, buf_b = 0;
  auto emit = [&](int32_t a, int32_t b) {
    if (np < cap) {
      const bool mine = lane == static_cast<int>(np & 63u);
      buf_a = mine ? a : buf_a;
      buf_b = mine ? b : buf_b;
      if ((np & 63u) == 63u) {
        const uint32_t k = np - 63u + static_cast<uint32_t>(lane);
        *reinterpret_cast<int2*>(out + 2ull * k) = make_int2(buf_a, buf_b);
      }
    } else {
      status = kDevCapacity;
    }
    ++np;
  };
  int32_t i = static_cast<int32_t>(Uniform(static_cast<uint32_t>(max_i)));
  int32_t j = static_cast<int32_t>(Uniform(static_cast<uint32_t>(max_j)));
  int32_t prev_i = 0, prev_j = 0;
  // the scalar engine's global walk ends only at the corner
  const bool corner_walk = sisd && type == 1;
  while (true) {
    // every step emits one pair, so a full buffer also bounds the walk
    if (status != kDevOk) break;
    if (corner_walk ? (i == 0 && j == -1) : (j == -1 || i == 0)) break;
    if (corner_walk && i == 0) {
      // Row 0 of the scalar walk: its cells hold the gap scores of the
      // read's prefix, so every step (a gap opened or extended, one column
      // each) moves left until column -1, as the SIMD engine's tail does.
      while (j != -1 && status == kDevOk) {
        emit(-1, j);
        --j;
      }
      break;
    }
    ensure(i, j);
    // Diagonal run through first predecessors, all in the tile (the
    // diagonal test below with k = 0 and j >= 1): each step reads the
    // predecessor's cell, its record and the next base in one round.
    if (T.on && i >= T.rb && i <= T.r1 && j >= T.c0 + 1 && j < T.c0 + T.C) {
      const unsigned long long td0 = SPOA_PH_NOW();
      (void)td0;
      bool moved = false;
      int32_t h = at(0, i, j);
      U32x4 rc = T.rec[i - T.rb];
      int32_t p0 = U(static_cast<int32_t>(rc.x));
      while (status == kDevOk && !(type == 0 && h == 0) && p0 >= T.r0 && j >= T.c0 + 1 &&
             j < static_cast<int32_t>(n)) {
        // the three reads are independent: issue them together
        const int32_t hd_v = static_cast<int32_t>(T.m[(((p0 - T.r0) * NM) << T.cshift) + (j - 1 - T.c0)]);
        const uint32_t b_v = T.sq[j - T.c0];
        const U32x4 rn = T.rec[max(p0 - T.rb, 0)];  // used only when p0 >= rb
        // keeps the record's read next to the others (not sunk past the test)
        __asm__ volatile("" ::"v"(rn.x), "v"(rn.y), "v"(rn.z));
        const int32_t hd = U(hd_v);
        const int32_t pr = Uniform(b_v) == Uniform(rc.y) ? prm.m : prm.mis;
        if (h != hd + pr) break;
        emit(U(static_cast<int32_t>(rc.z)), j);
        i = p0;
        --j;
        h = hd;
        rc = rn;
        p0 = U(static_cast<int32_t>(rc.x));
        moved = true;
        SPOA_TB_ADD(kTbDiagSteps, 1);
        if (i == 0 || i < T.rb) break;
      }
      SPOA_TB_ADD(kTbDiag, SPOA_PH_NOW() - td0);
      if (moved) continue;
    }
    const unsigned long long tg0 = SPOA_PH_NOW();
    (void)tg0;
    // The general step: one lane a predecessor, 64 at a time; the first
    // hit in predecessor order wins, as in the serial loops.
    const int32_t h = cell(0, i, j);
    if (type == 0 && h == 0) break;
    // row 0 has no predecessors (the scalar walk along it moves left only)
    const bool row0 = SISD && i == 0;
    const uint32_t pb = row0 ? 0u : pred_begin(i), pe = row0 ? 0u : pred_end(i);
    const bool source = pb == pe;
    const uint32_t npr = row0 ? 0u : (source ? 1u : pe - pb);
    bool found = false, ext_left = false, ext_up = false;
    // diagonal (:1053-1063); none from the first column
    const bool col_m1 = SISD && j < 0;
    const int32_t pr = !col_m1 && !row0 ? prof(i, j) : 0;
    for (uint32_t k0 = 0; k0 < npr && !found && !col_m1; k0 += 64) {
      const uint32_t k = k0 + static_cast<uint32_t>(lane);
      const bool act = k < npr;
      const int32_t p = act && !source ? pred_v(pb + k) : 0;
      bool hit = false;
      if (act) hit = h == (j == 0 ? gv.fcH[p] : at_v(0, p, j - 1)) + pr;
      const uint64_t hm = __ballot(hit);
      if (hm != 0) {
        prev_i = __builtin_amdgcn_readlane(p, static_cast<int>(__builtin_ctzll(hm)));
        prev_j = j - 1;
        found = true;
      }
    }
    // up (:1065-1074, :1447-1457, :1963-1975); the gap extends where the
    // winning test was an extension
    for (uint32_t k0 = 0; k0 < npr && !found; k0 += 64) {
      const uint32_t k = k0 + static_cast<uint32_t>(lane);
      const bool act = k < npr;
      const int32_t p = act && !source ? pred_v(pb + k) : 0;
      bool hit = false, ext = false;
      if (act) {
        const int32_t hu = cell_v(0, p, j);
        if (SUB == 0) {
          hit = h == hu + g;
        } else if (SUB == 1) {
          const bool fh = h == cell_v(1, p, j) + e;
          hit = fh || h == hu + g;
          ext = fh;
        } else {
          const bool fh = h == cell_v(1, p, j) + e, gh = h == hu + g, oh = h == cell_v(3, p, j) + c,
                     qh = h == hu + q;
          hit = fh || gh || oh || qh;
          ext = fh || (!gh && oh);
        }
      }
      const uint64_t hm = __ballot(hit);
      if (hm != 0) {
        const int w = static_cast<int>(__builtin_ctzll(hm));
        prev_i = __builtin_amdgcn_readlane(p, w);
        ext_up = __builtin_amdgcn_readlane(ext ? 1 : 0, w) != 0;
        prev_j = j;
        found = true;
      }
    }
    // left (:1076-1083, :1459-1468, :1977-1990); column 0 reads the first
    // column (the SIMD engine's E / Q there are its H, the scalar engine's
    // -inf); none from the first column itself
    if (!found && !col_m1) {
      const int32_t hl = j > 0 ? at(0, i, j - 1) : U(gv.fcH[i]);
      bool hit;
      if (SUB == 0) {
        hit = h == hl + g;
      } else {
        const int32_t el = j > 0 ? at(2, i, j - 1) : (sisd ? cell(2, i, -1) : U(gv.fcH[i]));
        if (SUB == 1) {
          hit = (ext_left = (h == el + e)) || h == hl + g;
        } else {
          const int32_t ql = j > 0 ? at(4, i, j - 1) : (sisd ? cell(4, i, -1) : U(gv.fcH[i]));
          hit = (ext_left = (h == el + e)) || h == hl + g || (ext_left = (h == ql + c)) || h == hl + q;
        }
      }
      if (hit) {
        prev_i = i;
        prev_j = j - 1;
        found = true;
      }
    }
    emit(i == prev_i ? -1 : node_of(i), j == prev_j ? -1 : j);
    i = prev_i;
    j = prev_j;
    const unsigned long long tg1 = SPOA_PH_NOW();
    (void)tg1;
    SPOA_TB_ADD(kTbGeneral, tg1 - tg0);
    SPOA_TB_ADD(kTbGenerals, 1);

    if (ext_left) {
      // :1484-1509, :2006-2039 -- convex stops when either chain breaks
      // (the scalar engine: when both break, sisd...:880-887).  The walk
      // emits columns j, j - 1, ... and stops at the first column jj < j
      // (going left) that is -1 or where the chain breaks; the chain is
      // tested 64 columns at a time.
      int32_t jj = j - 1;
      while (true) {
        const int32_t col = jj - lane;
        bool brk = col < 0;
        if (!brk) {
          const bool eb = at_v(2, i, col) + e != at_v(2, i, col + 1);
          if (SUB == 2) {
            const bool qb = at_v(4, i, col) + c != at_v(4, i, col + 1);
            brk = sisd ? (eb && qb) : (eb || qb);
          } else {
            brk = eb;
          }
        }
        const uint64_t bm = __ballot(brk);
        if (bm != 0) {
          jj -= static_cast<int32_t>(__builtin_ctzll(bm));
          break;
        }
        jj -= 64;
      }
      for (int32_t col = j; col > jj; --col) {
        emit(-1, col);
        if (status != kDevOk) {
          jj = col - 1;
          break;
        }
      }
      j = jj;
      SPOA_TB_ADD(kTbLeft, SPOA_PH_NOW() - tg1);
    } else if (ext_up) {
      // :1510-1547 affine, :2040-2093 convex
      while (i != 0 && status == kDevOk) {
        ensure(i, j);
        const uint32_t qb = pred_begin(i), qe = pred_end(i);
        const uint32_t nq = qe - qb;
        const int32_t fv = cell(1, i, j);
        bool stop;
        prev_i = 0;
        if (SUB == 1) {
          stop = false;
          for (uint32_t k0 = 0; k0 < nq; k0 += 64) {
            const uint32_t k = k0 + static_cast<uint32_t>(lane);
            const int32_t p = k < nq ? pred_v(qb + k) : 0;
            const bool gh = k < nq && fv == cell_v(0, p, j) + g;
            const bool fh = k < nq && fv == cell_v(1, p, j) + e;
            const uint64_t hm = __ballot(gh || fh);
            if (hm != 0) {
              const int w = static_cast<int>(__builtin_ctzll(hm));
              prev_i = __builtin_amdgcn_readlane(p, w);
              stop = __builtin_amdgcn_readlane(gh ? 1 : 0, w) != 0;
              break;
            }
          }
        } else {
          const int32_t ov = cell(3, i, j);
          stop = true;
          for (uint32_t k0 = 0; k0 < nq && stop; k0 += 64) {
            const uint32_t k = k0 + static_cast<uint32_t>(lane);
            const int32_t p = k < nq ? pred_v(qb + k) : 0;
            const bool hit = k < nq && (fv == cell_v(1, p, j) + e || ov == cell_v(3, p, j) + c);
            const uint64_t hm = __ballot(hit);
            if (hm != 0) {
              prev_i = __builtin_amdgcn_readlane(p, static_cast<int>(__builtin_ctzll(hm)));
              stop = false;
            }
          }
          if (stop) {
            for (uint32_t k0 = 0; k0 < nq; k0 += 64) {
              const uint32_t k = k0 + static_cast<uint32_t>(lane);
              const int32_t p = k < nq ? pred_v(qb + k) : 0;
              bool hit = false;
              if (k < nq) {
                const int32_t hv = cell_v(0, p, j);
                hit = fv == hv + g || ov == hv + q;
              }
              const uint64_t hm = __ballot(hit);
              if (hm != 0) {
                prev_i = __builtin_amdgcn_readlane(p, static_cast<int>(__builtin_ctzll(hm)));
                break;
              }
            }
          }
        }
        emit(node_of(i), -1);
        i = prev_i;
        if (stop || i == 0) break;
      }
      SPOA_TB_ADD(kTbUp, SPOA_PH_NOW() - tg1);
    }
  }

  // NW: finish along row 0 / column 0 (:1102-1124, :1552-1574, :2098-2121);
  // the scalar engine's main loop walked them already
  if (type == 1 && !sisd) {
    while (i == 0 && j != -1 && status == kDevOk) {
      emit(-1, j);
      --j;
    }
    while (i != 0 && j == -1 && status == kDevOk) {
      emit(gv.node_id[i - 1], -1);
      const uint32_t qb = gv.pred_off[i - 1], qe = gv.pred_off[i];
      if (qb == qe) {
        i = 0;
        continue;
      }
      int32_t next = i;
      for (uint32_t k = qb; k < qe; ++k) {
        const int32_t p = static_cast<int32_t>(gv.preds[k]);
        bool hit;
        if (SUB == 0) {
          hit = gv.fcH[i] == gv.fcH[p] + g;
        } else if (SUB == 1) {
          hit = gv.fcH[i] == gv.fcH[p] + e;
        } else {
          hit = gv.fcF[i] == gv.fcF[p] + e || gv.fcO[i] == gv.fcO[p] + c;
        }
        if (hit) {
          next = p;
          break;
        }
      }
      if (next == i) {
        status = kDevInconsistent;
        break;
      }
      i = next;
    }
  }
  // the last, partial batch of pairs
  const uint32_t kept = min(np, cap);
  if ((kept & 63u) != 0u) {
    const uint32_t k = (kept & ~63u) + static_cast<uint32_t>(lane);
    if (k < kept) *reinterpret_cast<int2*>(out + 2ull * k) = make_int2(buf_a, buf_b);
  }
  *n_out = np;
  *status_out = status;
}

// ------------------------------------------------------------ word ops --
// A lane's K columns of a score row live in 32-bit words: two int16
// columns a word (column 2i in the low half) or one int32 column a word.
// The int16 words are worked on with packed 16-bit instructions, whose
// wrap-around is the reference's int16 SIMD arithmetic.
typedef short S16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ S16x2 AsPk(uint32_t x) { return __builtin_bit_cast(S16x2, x); }
__device__ __forceinline__ uint32_t AsU(S16x2 x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ uint32_t PkAdd(uint32_t a, uint32_t b) { return AsU(AsPk(a) + AsPk(b)); }
// saturating: a carry term far below every real score stays at -32768
// instead of wrapping round
__device__ __forceinline__ uint32_t PkAddSat(uint32_t a, uint32_t b) {
  return AsU(__builtin_elementwise_add_sat(AsPk(a), AsPk(b)));
}
__device__ __forceinline__ uint32_t PkMax(uint32_t a, uint32_t b) {
  return AsU(__builtin_elementwise_max(AsPk(a), AsPk(b)));
}
__device__ __forceinline__ uint32_t DupLo(uint32_t a) {
  const S16x2 v = AsPk(a);
  return AsU(S16x2{v.x, v.x});
}
__device__ __forceinline__ uint32_t DupHi(uint32_t a) {
  const S16x2 v = AsPk(a);
  return AsU(S16x2{v.y, v.y});
}
__device__ __forceinline__ uint32_t SwapHalves(uint32_t a) {
  const S16x2 v = AsPk(a);
  return AsU(S16x2{v.y, v.x});
}
__device__ __forceinline__ uint32_t Pack16(int32_t lo, int32_t hi) {
  return (static_cast<uint32_t>(lo) & 0xffffu) | (static_cast<uint32_t>(hi) << 16);
}
__device__ __forceinline__ int32_t Lo16(uint32_t w) { return static_cast<int32_t>(static_cast<int16_t>(w)); }
__device__ __forceinline__ int32_t Hi16(uint32_t w) { return static_cast<int32_t>(w) >> 16; }
__device__ __forceinline__ int32_t Clamp16(int32_t v) { return min(max(v, -32768), 32767); }
// column k of packed words, in both halves
template <int KW>
__device__ __forceinline__ uint32_t HDup(const uint32_t (&w)[KW], int k) {
  return (k & 1) ? DupHi(w[k >> 1]) : DupLo(w[k >> 1]);
}

template <typename ST>
struct Cols {
  static constexpr int kPerWord = sizeof(ST) == 2 ? 2 : 1;
};

template <typename ST>
__device__ __forceinline__ uint32_t Splat(int32_t v) {
  if constexpr (sizeof(ST) == 2) {
    return (static_cast<uint32_t>(v) & 0xffffu) * 0x10001u;
  } else {
    return static_cast<uint32_t>(v);
  }
}

// a + b per column (b already splat)
template <typename ST>
__device__ __forceinline__ uint32_t WAdd(uint32_t a, uint32_t b) {
  if constexpr (sizeof(ST) == 2) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(S16x2, a) + __builtin_bit_cast(S16x2, b));
  } else {
    return a + b;
  }
}

template <typename ST>
__device__ __forceinline__ uint32_t WMax(uint32_t a, uint32_t b) {
  if constexpr (sizeof(ST) == 2) {
    return __builtin_bit_cast(
        uint32_t, __builtin_elementwise_max(__builtin_bit_cast(S16x2, a), __builtin_bit_cast(S16x2, b)));
  } else {
    return static_cast<uint32_t>(max(static_cast<int32_t>(a), static_cast<int32_t>(b)));
  }
}

// column k (of the word array) as a 32-bit score
template <typename ST, int KW>
__device__ __forceinline__ int32_t Col(const uint32_t (&w)[KW], int k) {
  if constexpr (sizeof(ST) == 2) {
    return (k & 1) ? static_cast<int32_t>(w[k >> 1]) >> 16 : static_cast<int32_t>(static_cast<int16_t>(w[k >> 1]));
  } else {
    return static_cast<int32_t>(w[k]);
  }
}

// the last column of a word array
template <typename ST, int KW>
__device__ __forceinline__ int32_t LastCol(const uint32_t (&w)[KW]) {
  if constexpr (sizeof(ST) == 2) {
    return static_cast<int32_t>(w[KW - 1]) >> 16;
  } else {
    return static_cast<int32_t>(w[KW - 1]);
  }
}

// Buffer resource over one DP row (all its matrices); out-of-range
// accesses read 0 and drop stores.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t RowRsrc(const void* row, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(row), 0, static_cast<int>(bytes), 0x00020000);
}

template <int KW>
__device__ __forceinline__ void BufLoad(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t (&w)[KW]) {
  static_assert(KW % 4 == 0, "rows move in 16-byte accesses");
#pragma unroll
  for (int i = 0; i < KW / 4; ++i) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16u * i, 0, 0);
    w[4 * i + 0] = v[0];
    w[4 * i + 1] = v[1];
    w[4 * i + 2] = v[2];
    w[4 * i + 3] = v[3];
  }
}

template <int KW>
__device__ __forceinline__ void BufStore(__amdgpu_buffer_rsrc_t r, uint32_t off, const uint32_t (&w)[KW]) {
#pragma unroll
  for (int i = 0; i < KW / 4; ++i) {
    typedef uint32_t U4 __attribute__((ext_vector_type(4)));
    U4 v;
    v[0] = w[4 * i + 0];
    v[1] = w[4 * i + 1];
    v[2] = w[4 * i + 2];
    v[3] = w[4 * i + 3];
    __builtin_amdgcn_raw_buffer_store_b128(v, r, off + 16u * i, 0, 0);
  }
}

template <typename ST>
__device__ __forceinline__ int32_t BufLoadCol(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(ST) == 2) {
    return static_cast<int32_t>(static_cast<int16_t>(__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0)));
  } else {
    return static_cast<int32_t>(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  }
}

// Packed int16 profile words (2 columns a word) of a lane's K columns for
// the row base rch: m on a match, n on a mismatch, pad past the read.
template <int K>
__device__ __forceinline__ void ProfileWords(const uint32_t (&scp)[K / 4], int32_t nvalid, uint32_t rch, int32_t m,
                                             int32_t mis, int32_t pad, uint32_t (&pw)[K / 2]) {
#pragma unroll
  for (int i = 0; i < K / 2; ++i) {
    const uint32_t c0 = (scp[i / 2] >> (16 * (i % 2))) & 0xffu;
    const uint32_t c1 = (scp[i / 2] >> (16 * (i % 2) + 8)) & 0xffu;
    const int32_t p0 = 2 * i < nvalid ? (c0 == rch ? m : mis) : pad;
    const int32_t p1 = 2 * i + 1 < nvalid ? (c1 == rch ? m : mis) : pad;
    pw[i] = (static_cast<uint32_t>(p0) & 0xffffu) | (static_cast<uint32_t>(p1) << 16);
  }
}

// --------------------------------------------------------------- kernel --
// SUB: 0 linear, 1 affine, 2 convex.  ST: stored score type.  K: columns
// per lane.  MAXT: launch bound (threads per workgroup).
//
// Dynamic LDS: the job's query profile, 4 row bases x K/2 words per lane
// (8 * K * blockDim.x bytes, each lane's words contiguous), used when a row
// is one tile and the graph has at most 4 distinct bases (the profile is
// otherwise formed per row).
// Occupancy target: three waves a SIMD (at most 168 registers a lane) where
// the row state fits with little spilling -- more waves in flight hide more
// of the per-row memory latency; wider lanes keep all their registers.
template <typename ST, int K>
struct WavesPerSimd {
  static constexpr int value = (sizeof(ST) == 2 ? K <= 16 : K <= 8) ? 3 : 1;
};

// L: the reference engine's lane count, whose row width and convex E/Q
// segments the launch reproduces -- 16 (its int16 engine) or 8 (its int32
// engine).  Int16 rows with L = 8 hold convex local alignments for which
// the reference picks int32 scores (its bound covers any alignment type)
// although a local alignment's scores stay within int16: every H is in
// [0, m * n] and every E, F, O, Q is at least its opening penalty above a
// non-negative H (the host admits a job only when m * (n + 8) leaves int16
// headroom).
// SPLIT: the launch may hold split jobs (row tiles on workgroups of their
// own, handing boundaries over through device memory); without it the row
// loop carries none of the hand-off code.  SISD: the launch reproduces the
// reference's scalar engine (its end cell, clamped local gap openings and
// traceback walk).  Both are per launch, so each is a kernel instance
// rather than a test in every row.
template <int SUB, typename ST, int K, int MAXT, int L, bool SPLIT, bool SISD>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WavesPerSimd<ST, K>::value)))
PoaAlignKernel(const KArgs a) {
  constexpr int CPW = Cols<ST>::kPerWord;
  constexpr int KW = K / CPW;  // score words a lane
  constexpr int PW = K / 2;    // profile words a lane
  constexpr int S = sizeof(ST);
  static_assert(SUB != 2 || K % L == 0, "convex lanes must start reference segments");
  static_assert(KW % 4 == 0 && PW % 4 == 0, "rows move in 16-byte accesses");
  constexpr int32_t kMinI = -2147483647 - 1;
#ifdef SPOA_ROW_PHASES
  unsigned long long ph[kPhCount] = {};
#endif

  __shared__ int32_t s_tot[2][kMaxWaves][2];
  __shared__ int32_t s_bnd[kMaxWaves];
  __shared__ int32_t s_red[kMaxWaves][3];
  __shared__ uint32_t s_job;
  extern __shared__ uint32_t s_prof[];

  const KParams& prm = a.prm;
  const int tid = threadIdx.x;
  const int nthreads = blockDim.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwaves = nthreads >> 6;
  const int type = prm.type;
  constexpr bool sisd = SISD;
  const int32_t m = prm.m, mis = prm.mis, g = prm.g, e = prm.e, q = prm.q, cc = prm.c;
  const int32_t pad = prm.pad;
  const uint32_t G2 = Splat<ST>(g), E2 = Splat<ST>(e), Q2 = Splat<ST>(q), C2 = Splat<ST>(cc);
  // scan constants: the transfer across the lanes of this lane's row up to
  // it (the carry out of the previous row of 16 lanes reaches it), and, for
  // multi-wave rows, from lane 0 to it (the wave carry); the convex 2x2
  // matrices live in LDS (read once per row) to spare registers
  int32_t lda = 0, ldc = 0;
  __shared__ int4 s_lpa[64], s_lpc[64];
  if constexpr (SUB != 2) {
    lda = ((lane & 15) + 1) * prm.decay;
    if constexpr (MAXT > 64) ldc = lane * prm.decay;
  } else if (wave == 0) {
    int32_t m4[4];
    LanePow(prm, (lane & 15) + 1, m4);
    s_lpa[lane] = make_int4(m4[0], m4[1], m4[2], m4[3]);
    if constexpr (MAXT > 64) {
      LanePow(prm, lane >= 1 ? lane : 1, m4);
      s_lpc[lane] = make_int4(m4[0], m4[1], m4[2], m4[3]);
    }
  }

  const uint32_t n_items = a.items ? a.n_items : a.n_jobs;
  for (;;) {
    if (tid == 0) s_job = atomicAdd(a.queue, 1u);
    __syncthreads();
    const uint32_t ix = Uniform(s_job);
    __syncthreads();
    if (ix >= n_items) break;
    // the job and, for a split job, the row tile this workgroup aligns
    uint32_t jx = ix, tile = 0;
    if (SPLIT && a.items) {  // (a launch without split jobs: item ix is (ix, 0))
      const uint2 it = a.items[ix];
      jx = Uniform(it.x);
      tile = Uniform(it.y);
    }
    // The launch lasts as long as its largest jobs (the queue hands them
    // out first): their waves take the SIMDs' issue slots before those of
    // the smaller jobs beside them, which have time to spare.
    if (jx < a.prio_jobs) {
      __builtin_amdgcn_s_setprio(2);
    } else {
      __builtin_amdgcn_s_setprio(0);
    }

    // the job record, every field uniform (scalar registers)
    const DevJob Jv = a.jobs[jx];
    DevJob J;
    J.mat_addr = (static_cast<uint64_t>(Uniform(static_cast<uint32_t>(Jv.mat_addr >> 32))) << 32) |
                 Uniform(static_cast<uint32_t>(Jv.mat_addr));
    J.graph_off = Uniform(Jv.graph_off);
    J.seq_off = Uniform(Jv.seq_off);
    J.V = Uniform(Jv.V);
    J.n = Uniform(Jv.n);
    J.W = Uniform(Jv.W);
    J.stride = Uniform(Jv.stride);
    J.pair_off = Uniform(Jv.pair_off);
    J.cap = Uniform(Jv.cap);
    J.nslots = Uniform(Jv.nslots);
    J.slot_chars = Uniform(Jv.slot_chars);
    J.split = Uniform(Jv.split);
    J.xfer_off = Uniform(Jv.xfer_off);
    const uint32_t V = J.V, n = J.n, W = J.W, stride = J.stride;
    const GraphView gv = MakeView(a.graphs + J.graph_off, V);
    const uint8_t* __restrict__ seq = a.seqs + J.seq_off;
    ST* __restrict__ mat = reinterpret_cast<ST*>(J.mat_addr);
    const uint32_t NM = SUB == 0 ? 1u : (SUB == 1 ? 3u : 5u);
    const uint32_t rbytes = stride * NM * S;  // one DP row, all matrices
    const uint32_t mbytes = stride * S;       // one matrix of a row
    const uint32_t tile_w = static_cast<uint32_t>(nthreads) * K;
    const uint32_t ntiles = (stride + tile_w - 1) / tile_w;
    // The tiles of every row this workgroup aligns: all of them in turn, or
    // -- a split job, whose ntiles tiles run on as many workgroups -- one.
    // A split tile t > 0 receives, per row, the H of tile t - 1's last
    // column and the carry entering its first column through hand-off words
    // (xin); a tile before the last sends them (xout).
    const bool split = SPLIT && J.split > 1;
    const uint32_t t_begin = split ? tile : 0u, t_end = split ? tile + 1 : ntiles;
    const bool single = t_end - t_begin == 1;  // one tile a row: the previous row stays in registers
    unsigned long long* xin = nullptr;
    unsigned long long* xout = nullptr;
    if (split) {
      unsigned long long* xj = a.xfer + J.xfer_off;
      if (tile > 0) xin = xj + 3ull * (tile - 1) * (V + 1);
      if (tile + 1 < ntiles) xout = xj + 3ull * tile * (V + 1);
    }
    constexpr int kXw = SUB == 2 ? 3 : 2;  // hand-off words a row: last H, carry (two for convex)
    bool aborted = false;
    // A row is nseg = ntiles * nwaves wave segments.  The first lane of
    // segment sg >= 1 needs, for the diagonal, the H value left of it in
    // every predecessor row: it keeps them in bnd[sg][row] (its own stores
    // and loads), received at the end of each row through LDS or, within
    // one wave, from lane 63.  No wave reads another wave's global stores
    // while the rows run, so the row barriers only order LDS.
    const uint32_t nseg = ntiles * static_cast<uint32_t>(nwaves);
    int32_t* __restrict__ bnd =
        reinterpret_cast<int32_t*>(reinterpret_cast<uint8_t*>(mat) + static_cast<uint64_t>(V + 1) * rbytes);
    const uint32_t nslots = single ? J.nslots : 0u;
    auto row_ptr = [&](uint32_t r) -> ST* { return mat + static_cast<uint64_t>(r) * stride * NM; };

    // row 0 (:556-694): E = g + c*e, Q = q + c*c, F = O = -inf, H per type
    auto row0_h = [&](int32_t ci) -> int32_t {
      if (type == 0) return 0;
      if (SUB == 0) return g + ci * g;
      if (SUB == 1) return g + ci * e;
      return max(g + ci * e, q + ci * cc);
    };
    const uint32_t col_end = min(stride, t_end * tile_w);
    for (uint32_t col = t_begin * tile_w + tid; col < col_end; col += nthreads) {
      const int32_t ci = static_cast<int32_t>(col);
      mat[col] = static_cast<ST>(row0_h(ci));
      if (SUB >= 1) {
        mat[stride + col] = static_cast<ST>(prm.ninf);
        mat[2 * stride + col] = static_cast<ST>(g + ci * e);
      }
      if (SUB == 2) {
        mat[3 * stride + col] = static_cast<ST>(prm.ninf);
        mat[4 * stride + col] = static_cast<ST>(q + ci * cc);
      }
    }

    // the read's bases under this lane's columns (reloaded per tile when a
    // row spans several tiles) and, with few distinct row bases, the
    // lane's query profile in LDS (written and read by this lane only)
    uint32_t scp[K / 4];
    int32_t nvalid;
    const uint32_t lane_c0 = t_begin * tile_w + static_cast<uint32_t>(tid) * K;  // first tile's columns
    LoadChars<K>(seq, lane_c0, n, lane_c0 < stride, scp, nvalid);
    for (uint32_t sl = 0; sl < nslots; ++sl) {
      uint32_t pw[PW];
      ProfileWords<K>(scp, nvalid, (J.slot_chars >> (8 * sl)) & 0xffu, m, mis, pad, pw);
#pragma unroll
      for (int i = 0; i < PW; ++i) s_prof[(sl * nthreads + tid) * PW + i] = pw[i];
    }

    // the previous row (H, F, O words of this lane's columns), from row 0
    uint32_t cH[KW], cF[KW], cO[KW];
    if (single) {
#pragma unroll
      for (int i = 0; i < KW; ++i) {
        const int32_t c = static_cast<int32_t>(lane_c0) + i * CPW;
        if constexpr (CPW == 2) {
          cH[i] = (static_cast<uint32_t>(row0_h(c)) & 0xffffu) | (static_cast<uint32_t>(row0_h(c + 1)) << 16);
        } else {
          cH[i] = static_cast<uint32_t>(row0_h(c));
        }
        cF[i] = Splat<ST>(prm.ninf);
        cO[i] = Splat<ST>(prm.ninf);
      }
    }
    __syncthreads();

    // End-cell candidates per lane: the lane's best (floored) value and the
    // first row reaching it.  SW starts at 0, OV and NW at the type's -inf.
    // The scalar engine starts OV and NW at its int32 -inf (sisd...:13-14).
    int32_t best = type == 0 ? 0 : (sisd ? kMinI + 1024 : prm.ninf);
    int32_t best_row = -1;
    int par = 0;
    int32_t fch_prev = gv.fcH[0];
    // row records two ahead, and the first predecessor row of the next row
    // when it is not the current one: loaded before the current row's
    // stores, so waiting for it never waits for those stores
    uint4 rec = LoadRow(gv, 0);
    uint4 rec2 = V > 1 ? LoadRow(gv, 1) : make_uint4(0u, 0u, 0u, 0u);
    uint32_t pfH[KW], pfF[KW], pfO[KW];
    bool pf_valid = false;
    uint32_t pf_row = 0;
    int32_t pf_hm1 = 0;  // with the prefetch: the predecessor's H left of the segment (first lane)
    // one-tile rows: H of the previous row left of this wave's segment
    int32_t bprev = 0;
    const uint32_t seg_c0 = t_begin * tile_w + static_cast<uint32_t>(wave) * 64u * K;
    if (single && seg_c0 > 0) {
      bprev = static_cast<int32_t>(static_cast<ST>(row0_h(static_cast<int32_t>(seg_c0) - 1)));
    }

    for (uint32_t r = 1; r <= V; ++r) {
      // row record: base | sink << 8 | profile slot << 9 | predecessor count
      // << 12, first predecessor row, offset of the predecessor list,
      // first-column score
      const uint4 cur = rec;
      const uint4 nxt = rec2;  // row r + 1 (when r < V)
      rec = rec2;
      if (r + 1 < V) rec2 = LoadRow(gv, r + 1);
      const uint32_t rch = cur.x & 0xffu;
      const bool sink = ((cur.x >> 8) & 1u) != 0;
      const uint32_t slot = (cur.x >> 9) & 7u;
      const uint32_t npred = cur.x >> 12;
      const uint32_t p0 = cur.y, pb = cur.z;
      const int32_t fch_r = static_cast<int32_t>(cur.w);
      ST* __restrict__ row = row_ptr(r);
      const __amdgpu_buffer_rsrc_t rrow = RowRsrc(row, rbytes);
      // carry entering column 0: linear H, affine E = fc + g, convex (E, Q)
      int32_t y0 = fch_r + g;
      int32_t y1 = fch_r + q;

      // a split tile's hand-off words of this row, loaded now and checked
      // before the scan (the tile before is usually rows ahead)
      unsigned long long xw = 0ull;
      if (xin != nullptr && lane < kXw) xw = XferGet(xin + 3ull * r + lane);
      int32_t hl = 0;  // tile t - 1's H at its last column, this row

      for (uint32_t t = t_begin; t < t_end; ++t) {
        const uint32_t c0 = t * tile_w + static_cast<uint32_t>(tid) * K;
        const bool active = c0 < stride;
        const uint32_t off = c0 * S;  // byte offset of the lane's columns in a matrix
        const unsigned long long tm0 = SPOA_PH_NOW();

        // ---- query profile of the row base
        if (!single) LoadChars<K>(seq, c0, n, active, scp, nvalid);
        uint32_t pw[PW];
        if (slot < nslots) {
          const uint4* src = reinterpret_cast<const uint4*>(s_prof + (slot * nthreads + tid) * PW);
#pragma unroll
          for (int i = 0; i < PW / 4; ++i) {
            const uint4 v = src[i];
            pw[4 * i + 0] = v.x;
            pw[4 * i + 1] = v.y;
            pw[4 * i + 2] = v.z;
            pw[4 * i + 3] = v.w;
          }
        } else {
          ProfileWords<K>(scp, nvalid, rch, m, mis, pad, pw);
        }

        const unsigned long long tm1 = SPOA_PH_NOW();
        // ---- 1. vertical pass over the predecessors, in in-edge order
        uint32_t h1[KW], fv[KW], ov[KW];
        for (uint32_t pi = 0; pi < npred; ++pi) {
          const uint32_t p = pi == 0 ? p0 : Uniform(gv.preds[pb + pi]);
          const bool cached = single && p + 1 == r;
          const bool fetched = !cached && pi == 0 && pf_valid && pf_row == p;
          const __amdgpu_buffer_rsrc_t rp = RowRsrc(row_ptr(p), rbytes);
          // the predecessor's H (and F, O): kept in registers, prefetched,
          // or read now -- all reads of a row together, then waited for
          uint32_t hv[KW], fp[KW], op[KW];
          int32_t hb = 0;  // lane 0: H_p left of the segment, read from bnd
          if (cached) {
#pragma unroll
            for (int i = 0; i < KW; ++i) {
              hv[i] = cH[i];
              if (SUB >= 1) fp[i] = cF[i];
              if (SUB == 2) op[i] = cO[i];
            }
          } else if (fetched) {
#pragma unroll
            for (int i = 0; i < KW; ++i) {
              hv[i] = pfH[i];
              if (SUB >= 1) fp[i] = pfF[i];
              if (SUB == 2) op[i] = pfO[i];
            }
          } else {
            BufLoad<KW>(rp, active ? off : rbytes, hv);
            if (SUB >= 1) BufLoad<KW>(rp, active ? off + mbytes : rbytes, fp);
            if (SUB == 2) BufLoad<KW>(rp, active ? off + 3 * mbytes : rbytes, op);
            if (lane == 0 && c0 != 0 && p != 0) {
              hb = bnd[static_cast<uint64_t>(t * static_cast<uint32_t>(nwaves) + wave) * (V + 1) + p];
            }
            const unsigned long long tw0 = SPOA_PH_NOW();
            WaitVmem();
            SPOA_PH_ADD(kPhPredWait, SPOA_PH_NOW() - tw0);
          }
          // H_p[c0 - 1]: neighbouring lane, or the first column / memory
          // across waves and tiles
          int32_t hm1 = __shfl_up(LastCol<ST, KW>(hv), 1, 64);
          if (lane == 0) {
            if (c0 == 0) {
              hm1 = p + 1 == r ? fch_prev : gv.fcH[p];
            } else if (p == 0) {
              hm1 = static_cast<int32_t>(static_cast<ST>(row0_h(static_cast<int32_t>(c0) - 1)));
            } else if (single && p + 1 == r) {
              hm1 = bprev;
            } else if (fetched) {
              hm1 = pf_hm1;
            } else {
              hm1 = hb;
            }
          }
          // diagonal (:838-850): H_p[c - 1] + profile
          uint32_t dg[KW];
          if constexpr (CPW == 2) {
            dg[0] = WAdd<ST>((hv[0] << 16) | (static_cast<uint32_t>(hm1) & 0xffffu), pw[0]);
#pragma unroll
            for (int i = 1; i < KW; ++i) dg[i] = WAdd<ST>(__builtin_amdgcn_alignbit(hv[i], hv[i - 1], 16), pw[i]);
          } else {
#pragma unroll
            for (int i = 0; i < KW; ++i) {
              const int32_t pr = (i & 1) ? static_cast<int32_t>(pw[i >> 1]) >> 16
                                         : static_cast<int32_t>(static_cast<int16_t>(pw[i >> 1]));
              dg[i] = static_cast<uint32_t>((i == 0 ? hm1 : static_cast<int32_t>(hv[i - 1])) + pr);
            }
          }
          // vertical moves: linear H_p + g; affine / convex the gap rows
          uint32_t vf[KW], vo[KW];
          if (SUB == 0) {
#pragma unroll
            for (int i = 0; i < KW; ++i) vf[i] = WAdd<ST>(hv[i], G2);
          } else {
#pragma unroll
            for (int i = 0; i < KW; ++i) vf[i] = WMax<ST>(WAdd<ST>(hv[i], G2), WAdd<ST>(fp[i], E2));
            if (SUB == 2) {
#pragma unroll
              for (int i = 0; i < KW; ++i) vo[i] = WMax<ST>(WAdd<ST>(hv[i], Q2), WAdd<ST>(op[i], C2));
            }
          }
          if (pi == 0) {
#pragma unroll
            for (int i = 0; i < KW; ++i) {
              h1[i] = dg[i];
              fv[i] = vf[i];
              if (SUB == 2) ov[i] = vo[i];
            }
            // the prefetch registers are free once the first predecessor is
            // folded: the next prefetch leaves now, with the rest of this
            // row's work to land before the drain ahead of its stores
            // prefetch the first predecessor of the next (row, tile): the next
            // tile of this row, or the first tile of the next row unless that
            // predecessor is this row kept in registers.  Every predecessor
            // tile was written earlier by these same lanes.
            pf_valid = false;
            {
              bool want = false;
              uint32_t q = 0, qc0 = 0;
              if (t + 1 < t_end) {
                want = true;
                q = p0;
                qc0 = c0 + tile_w;
              } else if (r < V && !(single && nxt.y == r)) {
                want = true;
                q = nxt.y;
                qc0 = lane_c0;
              }
              if (want) {
                const __amdgpu_buffer_rsrc_t rq = RowRsrc(row_ptr(q), rbytes);
                const bool qa = qc0 < stride;
                const uint32_t qoff = qc0 * S;
                BufLoad<KW>(rq, qa ? qoff : rbytes, pfH);
                if (SUB >= 1) BufLoad<KW>(rq, qa ? qoff + mbytes : rbytes, pfF);
                if (SUB == 2) BufLoad<KW>(rq, qa ? qoff + 3 * mbytes : rbytes, pfO);
                if (lane == 0 && qc0 > 0 && q > 0) {
                  const uint32_t qt = t + 1 < t_end ? t + 1 : t_begin;
                  pf_hm1 = bnd[static_cast<uint64_t>(qt * static_cast<uint32_t>(nwaves) + wave) * (V + 1) + q];
                }
                pf_valid = true;
                pf_row = q;
              }
            }
          } else {
#pragma unroll
            for (int i = 0; i < KW; ++i) {
              h1[i] = WMax<ST>(h1[i], dg[i]);
              fv[i] = WMax<ST>(fv[i], vf[i]);
              if (SUB == 2) ov[i] = WMax<ST>(ov[i], vo[i]);
            }
          }
        }
        // H1: the best of the vertical and diagonal moves; F and O are final
#pragma unroll
        for (int i = 0; i < KW; ++i) {
          h1[i] = WMax<ST>(h1[i], fv[i]);
          if (SUB == 2) h1[i] = WMax<ST>(h1[i], ov[i]);
        }
        // The scalar engine opens horizontal gaps from the clamped local H
        // (sisd...:530-533, :752-767); clamping H1 first gives the same H
        // and those gap scores.
        if (sisd && type == 0) {
#pragma unroll
          for (int i = 0; i < KW; ++i) h1[i] = WMax<ST>(h1[i], 0u);
        }
        if (single) {
#pragma unroll
          for (int i = 0; i < KW; ++i) {
            cF[i] = fv[i];
            if (SUB == 2) cO[i] = ov[i];
          }
        }

        const unsigned long long tm2 = SPOA_PH_NOW();
        // ---- 2. local pass: the horizontal gaps inside the lane with the
        // carry absent.  loc holds, per column, the best horizontal value
        // reachable from inside the lane (linear H, affine E, convex (E, Q)),
        // and (b0, b1) what the lane hands its right neighbour.  Column c0
        // itself takes its value from the carry alone (loc[0] is unused
        // there for affine / convex).
        constexpr int NLOC = (SUB == 2 && CPW == 2) ? K : KW;
        uint32_t loc[NLOC];
        uint32_t qloc[(SUB == 2 && CPW == 1) ? K : 1];
        int32_t b0 = 0, b1 = 0;
        if constexpr (CPW == 2) {
          if constexpr (SUB == 0) {
            // word i = columns (2i, 2i+1): (a, max(b, a + g)), then the chain
            const uint32_t k0g = Pack16(0, g), kgg = Pack16(g, 2 * g);
#pragma unroll
            for (int i = 0; i < KW; ++i) {
              const uint32_t wp = PkMax(h1[i], PkAdd(DupLo(h1[i]), k0g));
              loc[i] = i == 0 ? wp : PkMax(wp, PkAdd(DupHi(loc[i - 1]), kgg));
            }
            b0 = Hi16(loc[KW - 1]) + g;
          } else if constexpr (SUB == 1) {
            // E[c] = max(H1[c-1] + g, E[c-1] + e) on (c-1, c) pairs
            const uint32_t k0e = Pack16(0, e), kee = Pack16(e, 2 * e);
#pragma unroll
            for (int i = 0; i < KW; ++i) {
              const uint32_t sh = i == 0 ? (h1[0] << 16) : __builtin_amdgcn_alignbit(h1[i], h1[i - 1], 16);
              const uint32_t gi = PkAdd(sh, G2);  // (H1[2i-1] + g, H1[2i] + g)
              if (i == 0) {
                loc[0] = gi;  // high half: E at column c0 + 1
              } else {
                const uint32_t a2 = PkMax(gi, PkAdd(DupLo(gi), k0e));
                loc[i] = PkMax(a2, PkAdd(DupHi(loc[i - 1]), kee));
              }
            }
            b0 = max(Hi16(h1[KW - 1]) + g, Hi16(loc[KW - 1]) + e);
          } else {
            // (E, Q) of every column in one word; a gap opens from H1
            // inside an L-segment and from the final H at a segment start
            const uint32_t GQ = Pack16(g, q), EC = Pack16(e, cc);
            loc[0] = 0u;
            loc[1] = PkAdd(HDup(h1, 0), GQ);
#pragma unroll
            for (int k = 2; k < K; ++k) {
              if (k % L == 0) {
                const uint32_t hf = PkMax(PkMax(loc[k - 1], SwapHalves(loc[k - 1])), HDup(h1, k - 1));
                loc[k] = PkMax(PkAdd(hf, GQ), PkAdd(loc[k - 1], EC));
              } else {
                loc[k] = PkMax(PkAdd(HDup(h1, k - 1), GQ), PkAdd(loc[k - 1], EC));
              }
            }
            const uint32_t hf = PkMax(PkMax(loc[K - 1], SwapHalves(loc[K - 1])), HDup(h1, K - 1));
            const uint32_t nb = PkMax(PkAdd(hf, GQ), PkAdd(loc[K - 1], EC));
            b0 = Lo16(nb);
            b1 = Hi16(nb);
          }
        } else {
          auto hk = [&](int k) -> int32_t { return static_cast<int32_t>(h1[k]); };
          if constexpr (SUB == 0) {
            int32_t x = hk(0);
            loc[0] = static_cast<uint32_t>(x);
#pragma unroll
            for (int k = 1; k < K; ++k) {
              x = max(hk(k), x + g);
              loc[k] = static_cast<uint32_t>(x);
            }
            b0 = x + g;
          } else if constexpr (SUB == 1) {
            int32_t x = hk(0) + g;
            loc[0] = 0u;
            loc[1] = static_cast<uint32_t>(x);
#pragma unroll
            for (int k = 2; k < K; ++k) {
              x = max(hk(k - 1) + g, x + e);
              loc[k] = static_cast<uint32_t>(x);
            }
            b0 = max(hk(K - 1) + g, x + e);
          } else {
            int32_t xe = hk(0) + g, xq = hk(0) + q;
            loc[0] = 0u;
            qloc[0] = 0u;
            loc[1] = static_cast<uint32_t>(xe);
            qloc[1] = static_cast<uint32_t>(xq);
#pragma unroll
            for (int k = 2; k < K; ++k) {
              const int32_t hp = hk(k - 1);
              const int32_t xs = k % L == 0 ? max(hp, max(xe, xq)) : hp;
              const int32_t ne = max(xs + g, xe + e);
              const int32_t nq = max(xs + q, xq + cc);
              xe = ne;
              xq = nq;
              loc[k] = static_cast<uint32_t>(xe);
              qloc[k] = static_cast<uint32_t>(xq);
            }
            const int32_t hf = max(hk(K - 1), max(xe, xq));
            b0 = max(hf + g, xe + e);
            b1 = max(hf + q, xq + cc);
          }
        }

        // a split tile after the first: the carry entering its first column
        // and tile t - 1's H at its last column come from the hand-off words
        // (after a failed wait they are not awaited again: the job reports
        // the time-out)
        if (xin != nullptr) {
          if (!aborted && !XferWait(a.xfer, a.epoch, a.wait_ticks, xin + 3ull * r, kXw, lane, xw)) aborted = true;
          const int32_t lo = static_cast<int32_t>(static_cast<uint32_t>(xw));
          hl = __builtin_amdgcn_readlane(lo, 0);
          y0 = __builtin_amdgcn_readlane(lo, 1);
          if constexpr (SUB == 2) y1 = __builtin_amdgcn_readlane(lo, 2);
        }

        const unsigned long long tm3 = SPOA_PH_NOW();
        // ---- wavefront inclusive scan of the lane transfers.  Inside each
        // row of 16 lanes a Kogge-Stone scan on DPP row shifts, the sender
        // applying the uniform step transfer M^sh (scalar: sh times the lane
        // decay); across rows, the inclusive values at the ends of rows 0-2
        // (read into scalar registers) are chained with M^16 and moved into
        // the next row by each lane's own M^(lane % 16 + 1).  A one-wave row
        // folds the tile carry into lane 0 first, so the exclusive carry of
        // every lane is its left neighbour's inclusive value.
        const bool one_wave = nwaves == 1;
        int32_t s0 = b0, s1 = b1;
        if (one_wave && lane == 0) {
          if constexpr (SUB != 2) {
            s0 = max(s0, y0 + prm.decay);
          } else {
            int32_t u0 = y0, u1 = y1;
            MpApply(prm.mpow[0], u0, u1);
            s0 = max(s0, u0);
            s1 = max(s1, u1);
          }
        }
        if constexpr (SUB != 2) {
          const int32_t d = prm.decay;
          s0 = max(s0, Dpp<kRowShr1, 0xf>(s0, s0 + d));
          s0 = max(s0, Dpp<kRowShr2, 0xf>(s0, s0 + 2 * d));
          s0 = max(s0, Dpp<kRowShr4, 0xf>(s0, s0 + 4 * d));
          s0 = max(s0, Dpp<kRowShr8, 0xf>(s0, s0 + 8 * d));
          const int32_t r1 = __builtin_amdgcn_readlane(s0, 15);
          const int32_t r2 = max(__builtin_amdgcn_readlane(s0, 31), r1 + 16 * d);
          const int32_t r3 = max(__builtin_amdgcn_readlane(s0, 47), r2 + 16 * d);
          const int32_t rr = lane < 32 ? r1 : (lane < 48 ? r2 : r3);
          if (lane >= 16) s0 = max(s0, rr + lda);
        } else {
#define SPOA_ROW_STEP(CTRL, T)                                               \
  {                                                                          \
    const int32_t u0 = max(prm.mpow[T][0] + s0, prm.mpow[T][1] + s1);        \
    const int32_t u1 = max(prm.mpow[T][2] + s0, prm.mpow[T][3] + s1);        \
    s0 = max(s0, Dpp<CTRL, 0xf>(s0, u0));                                    \
    s1 = max(s1, Dpp<CTRL, 0xf>(s1, u1));                                    \
  }
          SPOA_ROW_STEP(kRowShr1, 0)
          SPOA_ROW_STEP(kRowShr2, 1)
          SPOA_ROW_STEP(kRowShr4, 2)
          SPOA_ROW_STEP(kRowShr8, 3)
#undef SPOA_ROW_STEP
          int32_t a0 = __builtin_amdgcn_readlane(s0, 15), a1 = __builtin_amdgcn_readlane(s1, 15);
          int32_t c0r = a0, c1r = a1;  // carry out of row 0
          MpApply(prm.mpow[4], a0, a1);
          a0 = max(a0, __builtin_amdgcn_readlane(s0, 31));
          a1 = max(a1, __builtin_amdgcn_readlane(s1, 31));
          int32_t c0s = a0, c1s = a1;  // out of rows 0-1
          MpApply(prm.mpow[4], a0, a1);
          a0 = max(a0, __builtin_amdgcn_readlane(s0, 47));
          a1 = max(a1, __builtin_amdgcn_readlane(s1, 47));  // out of rows 0-2
          const int32_t z0 = lane < 32 ? c0r : (lane < 48 ? c0s : a0);
          const int32_t z1 = lane < 32 ? c1r : (lane < 48 ? c1s : a1);
          if (lane >= 16) {
            const int4 lp = s_lpa[lane];
            const int32_t n0 = max(lp.x + z0, lp.y + z1), n1 = max(lp.z + z0, lp.w + z1);
            s0 = max(s0, n0);
            s1 = max(s1, n1);
          }
        }
        // carry into this wave: the tile carry folded with earlier waves
        int32_t w0 = y0, w1 = y1;
        int32_t tile_out0, tile_out1;
        if (!one_wave) {
          if (lane == 63) {
            s_tot[par][wave][0] = s0;
            s_tot[par][wave][1] = s1;
          }
          const unsigned long long tb0 = SPOA_PH_NOW();
          LdsBarrier();
          SPOA_PH_ADD(kPhScanBarrier, SPOA_PH_NOW() - tb0);
          int32_t f0 = y0, f1 = y1;
          for (int w = 0; w < nwaves; ++w) {
            if (w == wave) {
              w0 = f0;
              w1 = f1;
            }
            const int32_t t0 = s_tot[par][w][0], t1 = s_tot[par][w][1];
            if (SUB != 2) {
              f0 = max(t0, f0 + 64 * prm.decay);
            } else {
              MpApply(prm.mpow[6], f0, f1);
              f0 = max(f0, t0);
              f1 = max(f1, t1);
            }
          }
          tile_out0 = f0;
          tile_out1 = f1;
          par ^= 1;
        } else {
          // lane 63 already holds the tile carry moved across the wave
          tile_out0 = __builtin_amdgcn_readlane(s0, 63);
          tile_out1 = SUB == 2 ? __builtin_amdgcn_readlane(s1, 63) : 0;
        }
        // exclusive carry into this lane
        int32_t x0 = Dpp<kWaveShr1, 0xf>(w0, s0), x1 = 0;
        if constexpr (SUB == 2) x1 = Dpp<kWaveShr1, 0xf>(w1, s1);
        if (!one_wave && lane > 0) {
          // the wave carry moved across the lanes before this one
          if constexpr (SUB != 2) {
            x0 = max(x0, w0 + ldc);
          } else {
            const int4 lc = s_lpc[lane];
            const int32_t n0 = max(lc.x + w0, lc.y + w1), n1 = max(lc.z + w0, lc.w + w1);
            x0 = max(x0, n0);
            x1 = max(x1, n1);
          }
        }

        const unsigned long long tm4 = SPOA_PH_NOW();
        // ---- 3. fix-up: every column's value is the better of its local
        // value and the carry moved to it; then H.  E / Q leave in 16-byte
        // chunks as soon as a chunk is complete, H after the row.  The
        // prefetch has had the row's work to land; it is waited for here,
        // before the row's first store (F, O), so that the next row never
        // waits for this row's stores.
        const unsigned long long ts0 = SPOA_PH_NOW();
        WaitVmem();
        SPOA_PH_ADD(kPhStoreWait, SPOA_PH_NOW() - ts0);
        if (SUB >= 1 && active) {
          BufStore<KW>(rrow, off + mbytes, fv);
          if (SUB == 2) BufStore<KW>(rrow, off + 3 * mbytes, ov);
        }
        uint32_t hw[KW];
        if constexpr (CPW == 2) {
          if constexpr (SUB == 0) {
            uint32_t v = PkAddSat(Splat<ST>(Clamp16(x0)), Pack16(0, g));
            const uint32_t k2g = Splat<ST>(2 * g);
#pragma unroll
            for (int i = 0; i < KW; ++i) {
              hw[i] = PkMax(loc[i], v);
              v = PkAddSat(v, k2g);
            }
          } else if constexpr (SUB == 1) {
            uint32_t v = PkAddSat(Splat<ST>(Clamp16(x0)), Pack16(0, e));
            const uint32_t k2e = Splat<ST>(2 * e);
#pragma unroll
            for (int j = 0; j < KW / 4; ++j) {
              uint32_t ew[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const int i = 4 * j + u;
                uint32_t x = PkMax(loc[i], v);
                if (i == 0) x = (x & 0xffff0000u) | (v & 0xffffu);  // E at c0 is the carry
                ew[u] = x;
                hw[i] = PkMax(h1[i], x);
                v = PkAddSat(v, k2e);
              }
              if (active) BufStore<4>(rrow, off + 2 * mbytes + 16u * j, ew);
            }
          } else {
            const uint32_t EC = Pack16(e, cc);
            int32_t cE = x0, cQ = x1;
            uint32_t v = 0u;
#pragma unroll
            for (int j = 0; j < KW / 4; ++j) {
              // columns 8j .. 8j + 7
#pragma unroll
              for (int t = 0; t < 8; ++t) {
                const int k = 8 * j + t;
                if (k % L == 0) {
                  if (k > 0) MpApply(prm.mseg, cE, cQ);
                  v = Pack16(Clamp16(cE), Clamp16(cQ));
                }
                loc[k] = k == 0 ? v : PkMax(loc[k], v);
                v = PkAddSat(v, EC);
              }
              uint32_t ew[4], qw[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const int i = 4 * j + u;
                ew[u] = __builtin_amdgcn_perm(loc[2 * i + 1], loc[2 * i], 0x05040100u);
                qw[u] = __builtin_amdgcn_perm(loc[2 * i + 1], loc[2 * i], 0x07060302u);
                hw[i] = PkMax(h1[i], PkMax(ew[u], qw[u]));
              }
              if (active) {
                BufStore<4>(rrow, off + 2 * mbytes + 16u * j, ew);
                BufStore<4>(rrow, off + 4 * mbytes + 16u * j, qw);
              }
            }
          }
          if (type == 0) {
#pragma unroll
            for (int i = 0; i < KW; ++i) hw[i] = PkMax(hw[i], 0u);
          }
        } else {
          if constexpr (SUB == 0) {
            int32_t v = x0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
              hw[k] = static_cast<uint32_t>(max(static_cast<int32_t>(loc[k]), v));
              v += g;
            }
          } else if constexpr (SUB == 1) {
            int32_t v = x0;
#pragma unroll
            for (int j = 0; j < K / 4; ++j) {
              uint32_t ew[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const int k = 4 * j + u;
                const int32_t ek = k == 0 ? v : max(static_cast<int32_t>(loc[k]), v);
                ew[u] = static_cast<uint32_t>(ek);
                hw[k] = static_cast<uint32_t>(max(static_cast<int32_t>(h1[k]), ek));
                v += e;
              }
              if (active) BufStore<4>(rrow, off + 2 * mbytes + 16u * j, ew);
            }
          } else {
            int32_t cE = x0, cQ = x1, vE = 0, vQ = 0;
#pragma unroll
            for (int j = 0; j < K / 4; ++j) {
              uint32_t ew[4], qw[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const int k = 4 * j + u;
                if (k % L == 0) {
                  if (k > 0) MpApply(prm.mseg, cE, cQ);
                  vE = cE;
                  vQ = cQ;
                }
                const int32_t ek = k == 0 ? vE : max(static_cast<int32_t>(loc[k]), vE);
                const int32_t qk = k == 0 ? vQ : max(static_cast<int32_t>(qloc[k]), vQ);
                ew[u] = static_cast<uint32_t>(ek);
                qw[u] = static_cast<uint32_t>(qk);
                hw[k] = static_cast<uint32_t>(max(static_cast<int32_t>(h1[k]), max(ek, qk)));
                vE += e;
                vQ += cc;
              }
              if (active) {
                BufStore<4>(rrow, off + 2 * mbytes + 16u * j, ew);
                BufStore<4>(rrow, off + 4 * mbytes + 16u * j, qw);
              }
            }
          }
          if (type == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) hw[k] = static_cast<uint32_t>(max(static_cast<int32_t>(hw[k]), 0));
          }
        }
        if (active) BufStore<KW>(rrow, off, hw);
        if (single) {
#pragma unroll
          for (int i = 0; i < KW; ++i) cH[i] = hw[i];
        }
        // a split tile before the last: hand this row's last H and the carry
        // out of the tile to the next tile (its columns are all inside the
        // row, so the last lane of the last wave holds them)
        if (xout != nullptr && wave == nwaves - 1) {
          const int32_t last = LastCol<ST, KW>(hw);
          if (lane == 63) {
            unsigned long long* xo = xout + 3ull * r;
            XferPut(xo, a.epoch, last);
            XferPut(xo + 1, a.epoch, tile_out0);
            if constexpr (SUB == 2) XferPut(xo + 2, a.epoch, tile_out1);
          }
        }

        const unsigned long long tm5 = SPOA_PH_NOW();
        // end-cell candidates (:903-928): SW all rows, OV sink rows (row
        // maxima floored at 0, :252-262), NW sink rows at column n - 1.
        // The scalar engine takes the first strictly better cell in
        // row-major order over the read's columns only, without the floor
        // (sisd...:296-311): a row's maximum over columns < n.
        if (active && sisd && (type == 0 || (type == 2 && sink))) {
          int32_t lm = kMinI;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            if (c0 + k < n) lm = max(lm, Col<ST, KW>(hw, k));
          }
          if (best < lm) {
            best = lm;
            best_row = static_cast<int32_t>(r);
          }
        } else if (active) {
          if (type == 0 || (type == 2 && sink)) {
            // whole words up to the reference width W (a multiple of 8)
            const int32_t wv = (static_cast<int32_t>(W) - static_cast<int32_t>(c0)) / CPW;
            uint32_t mw = Splat<ST>(0);
#pragma unroll
            for (int i = 0; i < KW; ++i) {
              if (i < wv) mw = WMax<ST>(mw, hw[i]);
            }
            int32_t lm;
            if constexpr (CPW == 2) {
              lm = max(static_cast<int32_t>(static_cast<int16_t>(mw)), static_cast<int32_t>(mw) >> 16);
            } else {
              lm = static_cast<int32_t>(mw);
            }
            if (best < lm) {
              best = lm;
              best_row = static_cast<int32_t>(r);
            }
          } else if (type == 1 && sink) {
            const uint32_t cn = n - 1;
            if (cn >= c0 && cn < c0 + K) {
              int32_t v = Col<ST, KW>(hw, 0);
#pragma unroll
              for (int k = 1; k < K; ++k) {
                if (c0 + k == cn) v = Col<ST, KW>(hw, k);
              }
              if (best < v) {
                best = v;
                best_row = static_cast<int32_t>(r);
              }
            }
          }
        }
        y0 = tile_out0;
        y1 = tile_out1;
        // hand the H value of this segment's last column to the next
        // segment's first lane
        if (nseg > 1) {
          const int32_t last = LastCol<ST, KW>(hw);
          int32_t got = 0;
          uint32_t to_sg = 0;
          bool have = false;
          if (nwaves > 1) {
            if (lane == 63) s_bnd[wave] = last;
            const unsigned long long tb1 = SPOA_PH_NOW();
            LdsBarrier();
            SPOA_PH_ADD(kPhBoundaryBarrier, SPOA_PH_NOW() - tb1);
            if (lane == 0) {
              if (wave > 0) {
                got = s_bnd[wave - 1];
                to_sg = t * static_cast<uint32_t>(nwaves) + wave;
                have = true;
              } else if (t + 1 < t_end) {
                got = s_bnd[nwaves - 1];
                to_sg = (t + 1) * static_cast<uint32_t>(nwaves);
                have = true;
              } else if (xin != nullptr) {
                got = hl;
                to_sg = t * static_cast<uint32_t>(nwaves);
                have = true;
              }
            }
          } else {
            got = __builtin_amdgcn_readlane(last, 63);
            if (lane == 0 && t + 1 < t_end) {
              to_sg = t + 1;
              have = true;
            } else if (lane == 0 && xin != nullptr) {
              got = hl;
              to_sg = t;
              have = true;
            }
          }
          if (have) {
            bnd[static_cast<uint64_t>(to_sg) * (V + 1) + r] = got;
            if (single) bprev = got;
          }
        }
#ifdef SPOA_ROW_PHASES
        {
          const unsigned long long tm6 = SPOA_PH_NOW();
          ph[kPhProfile] += tm1 - tm0;
          ph[kPhVertical] += tm2 - tm1;
          ph[kPhLocal] += tm3 - tm2;
          ph[kPhScan] += tm4 - tm3;
          ph[kPhFixup] += tm5 - tm4;
          ph[kPhEndBoundary] += tm6 - tm5;
          ph[kPhTiles] += 1;
        }
#else
        (void)tm0, (void)tm1, (void)tm2, (void)tm3, (void)tm4, (void)tm5;
#endif
      }
      fch_prev = fch_r;
    }
    const unsigned long long tr0 = SPOA_PH_NOW();
    // the traceback reads what every wave stored
    __threadfence_block();

    // ---- end cell: the best value, then the first row reaching it
    {
      int32_t v = best, rr = best_row < 0 ? 0x7fffffff : best_row;
      // lanes without a candidate row do not compete
      if (best_row < 0) v = kMinI;
#pragma unroll
      for (int st = 32; st >= 1; st >>= 1) {
        const int32_t ov2 = __shfl_xor(v, st, 64);
        const int32_t or2 = __shfl_xor(rr, st, 64);
        if (ov2 > v || (ov2 == v && or2 < rr)) {
          v = ov2;
          rr = or2;
        }
      }
      if (lane == 0) {
        s_red[wave][0] = v;
        s_red[wave][1] = rr;
      }
    }
    __syncthreads();
    int32_t max_score = kMinI, max_i = 0x7fffffff;
    for (int w = 0; w < nwaves; ++w) {
      const int32_t v = s_red[w][0], rr = s_red[w][1];
      if (v > max_score || (v == max_score && rr < max_i)) {
        max_score = v;
        max_i = rr;
      }
    }
    __syncthreads();

    // A split job: a tile before the last publishes its candidate once its
    // rows are visible device-wide (the barrier above has waited for every
    // wave's stores) and is done; the last tile -- it finishes after all
    // the others, whose last rows it waited for -- takes the best of all
    // the candidates (the first row on ties, as across waves), acquires the
    // other tiles' rows and walks the traceback.
    bool last_tile = true;
    if (split) {
      unsigned long long* xc = a.xfer + J.xfer_off + 3ull * (ntiles - 1) * (V + 1);
      if (tile + 1 < ntiles) {
        last_tile = false;
        if (tid == 0) {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
          XferPut(xc + 2 * tile, a.epoch, max_score);
          XferPut(xc + 2 * tile + 1, a.epoch, max_i);
        }
      } else {
        if (wave == 0) {
          for (uint32_t u0 = 0; u0 + 1 < ntiles; u0 += 32) {
            const uint32_t nu = min(32u, ntiles - 1 - u0);
            const int nw = static_cast<int>(2 * nu);
            unsigned long long w = 0ull;
            if (lane < nw) w = XferGet(xc + 2ull * u0 + lane);
            if (!aborted && !XferWait(a.xfer, a.epoch, a.wait_ticks, xc + 2ull * u0, nw, lane, w)) aborted = true;
            const int32_t lo = static_cast<int32_t>(static_cast<uint32_t>(w));
            int32_t v = __shfl(lo, (2 * lane) & 63, 64), rr = __shfl(lo, (2 * lane + 1) & 63, 64);
            if (static_cast<uint32_t>(lane) >= nu) {
              v = kMinI;
              rr = 0x7fffffff;
            }
#pragma unroll
            for (int st = 16; st >= 1; st >>= 1) {
              const int32_t ov2 = __shfl_xor(v, st, 64);
              const int32_t or2 = __shfl_xor(rr, st, 64);
              if (ov2 > v || (ov2 == v && or2 < rr)) {
                v = ov2;
                rr = or2;
              }
            }
            v = __builtin_amdgcn_readfirstlane(v);
            rr = __builtin_amdgcn_readfirstlane(rr);
            if (v > max_score || (v == max_score && rr < max_i)) {
              max_score = v;
              max_i = rr;
            }
          }
          if (__builtin_amdgcn_ballot_w64(XferTagged(XferGet(a.xfer), a.epoch)) != 0) aborted = true;
          if (lane == 0) {
            s_red[0][0] = max_score;
            s_red[0][1] = max_i;
            s_red[0][2] = aborted ? 1 : 0;
          }
        }
        __syncthreads();
        max_score = s_red[0][0];
        max_i = s_red[0][1];
        aborted = s_red[0][2] != 0;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __syncthreads();
      }
    }
    const bool found = max_i != 0x7fffffff;

    // max_j (:937-953): NW the last column; SW/OV the first column of row
    // max_i (padding included; the scalar engine: the read's columns)
    // holding the best value, -1 if none
    const uint64_t rstride = static_cast<uint64_t>(stride) * NM;
    int32_t max_j = static_cast<int32_t>(n) - 1;
    if (last_tile && !aborted && found && type != 1) {
      const ST* __restrict__ hrow = mat + static_cast<uint64_t>(max_i) * rstride;
      const uint32_t cols = sisd ? n : W;
      int32_t first = 0x7fffffff;
      for (uint32_t col = tid; col < cols; col += nthreads) {
        if (static_cast<int32_t>(hrow[col]) == max_score) {
          first = static_cast<int32_t>(col);
          break;
        }
      }
#pragma unroll
      for (int st = 32; st >= 1; st >>= 1) first = min(first, __shfl_xor(first, st, 64));
      if (lane == 0) s_red[wave][2] = first;
      __syncthreads();
      first = 0x7fffffff;
      for (int w = 0; w < nwaves; ++w) first = min(first, s_red[w][2]);
      max_j = first == 0x7fffffff ? -1 : first;
      __syncthreads();
    }

    const unsigned long long tr1 = SPOA_PH_NOW();
    SPOA_PH_ADD(kPhReduce, tr1 - tr0);
    (void)tr0, (void)tr1;
    // ---- traceback by wave 0 (the query profile's LDS is free now)
    if (last_tile && wave == 0) {
      int32_t* res = a.results + 4 * jx;
      if (aborted) {
        if (lane == 0) {
          res[0] = 0;
          res[1] = 0;
          res[2] = 0;
          res[3] = kDevTimeout;
        }
      } else if (!found) {
        if (lane == 0) {
          res[0] = 0;
          res[1] = 0;
          res[2] = 0;
          res[3] = kDevOk;
        }
      } else {
        uint32_t np = 0;
        int32_t status = kDevOk;
        const unsigned long long tt0 = SPOA_PH_NOW();
        unsigned long long tph[kTbCount] = {};
        Traceback<SUB, ST, SISD>(prm, gv, seq, n, mat, rstride, stride, max_i, max_j, a.pairs + 2ull * J.pair_off,
                           J.cap, &np, &status, (Lds<uint32_t>*)(s_prof),
                           a.lds_bytes, lane, tph);
#ifdef SPOA_ROW_PHASES
        if (lane == 0 && a.phases) {
          for (int k = 0; k < kTbCount; ++k) atomicAdd(a.phases + kPhCount + k, tph[k]);
        }
#endif
        SPOA_PH_ADD(kPhTraceback, SPOA_PH_NOW() - tt0);
        (void)tt0;
        if (lane == 0) {
          res[0] = static_cast<int32_t>(np);
          res[1] = max_score;
          res[2] = 1;
          res[3] = status;
        }
      }
    }
#ifdef SPOA_ROW_PHASES
    if (wave == 0) {
      ph[kPhJobs] += 1;
      if (lane == 0 && a.phases) {
        for (int k = 0; k < kPhCount; ++k) atomicAdd(a.phases + k, ph[k]);
      }
    }
    for (int k = 0; k < kPhCount; ++k) ph[k] = 0;
#endif
    __syncthreads();
  }
}

}  // namespace spoa_hip_dev
